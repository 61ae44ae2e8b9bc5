/**
 * Discrete-time autoregressive model: x <- theta*x + w at every unit of
 * time, observed with Gaussian error.
 */
model AR1 {
  param theta, sigma2, obs_sd
  noise w
  state x
  obs y

  sub parameter {
    theta ~ gaussian()
    sigma2 ~ uniform(0.1, 1.0)
  }

  sub initial {
    x ~ gaussian(0.0, 1.0)
  }

  sub transition(delta = 1.0) {
    w ~ gaussian(0.0, sqrt(sigma2))
    x <- theta*x + w
  }

  sub observation {
    y ~ gaussian(mean = x, std = obs_sd)
  }
}
