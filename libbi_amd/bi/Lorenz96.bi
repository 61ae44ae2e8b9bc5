/**
 * Lorenz-96 with additive noise at every observation interval of 0.05 and
 * Gaussian observations of the even components: the model of the
 * hand-written Lorenz-96 functor, at eight components (one lane holds a
 * particle's whole state).
 */
model Lorenz96 {
  const N = 8
  dim n(size = N, boundary = 'cyclic')
  dim m(size = N/2)

  param F, noise_sd, obs_sd, init_sd
  state x[n]
  noise w[n]
  obs y[m]

  sub initial {
    x[i] ~ gaussian(0.0, init_sd)
  }

  sub transition(delta = 0.05) {
    ode(alg = 'RK5(4)', h = 0.05, atoler = 1.0e-5, rtoler = 1.0e-5) {
      dx[i]/dt = (x[i+1] - x[i-2])*x[i-1] - x[i] + F
    }
    w ~ gaussian()
    x <- x + noise_sd*w
  }

  sub observation {
    y[j] ~ gaussian(x[2*j], obs_sd)
  }
}
