/**
 * The Lorenz-96 filter of the benchmark, at forty components, from a model
 * file: the generator splits each particle's state over a group of four
 * lanes, as the hand-written functor does, and exchanges the stencil's
 * neighbours between them.
 */
model Lorenz96d40 {
  const N = 40
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
