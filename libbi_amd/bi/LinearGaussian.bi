/**
 * Scalar linear-Gaussian model: exponential decay integrated by DOPRI5, then
 * additive Gaussian noise; its likelihood is the Kalman filter's.
 */
model LinearGaussian {
  param lam, noise_sd, obs_sd, m0, s0
  state x
  noise w
  obs y

  sub parameter {
    lam ~ gamma(2.0, 0.25)
    noise_sd ~ inverse_gamma(3.0, 1.0)
    obs_sd ~ log_normal(log(0.5), 0.2)
    m0 ~ gaussian(0.0, 1.0)
    s0 ~ uniform(0.5, 2.0)
  }

  sub initial {
    x ~ gaussian(m0, s0)
  }

  sub transition(delta = 1.0) {
    w ~ gaussian()
    ode(alg = 'RK5(4)', h = 0.1, atoler = 1.0e-5, rtoler = 1.0e-5) {
      dx/dt = -lam*x
    }
    x <- x + noise_sd*w
  }

  sub observation {
    y ~ gaussian(x, obs_sd)
  }
}
