/**
 * Susceptible-infectious-recovered epidemic with a log-normal perturbation
 * of the infection rate at every unit of time and negative-binomial counts
 * of the new infections over each interval.
 */
model SIR {
  const N = 1000.0   // population size

  param beta, gamma, k, sigma  // infection rate, recovery rate, overdispersion, rate noise
  state S, I, R, Z             // compartments; Z: infections since the last time unit
  noise w
  obs incidence

  sub parameter {
    beta ~ gamma(4.0, 0.15)
    gamma ~ beta(2.0, 6.0)
    k ~ exponential(0.1)
    sigma ~ uniform(0.0, 0.3)
  }

  sub initial {
    S <- N - 10.0
    I <- 10.0
    R <- 0.0
    Z <- 0.0
  }

  sub transition(delta = 1.0) {
    w ~ gaussian(0.0, sigma)
    Z <- 0.0
    ode(alg = 'RK4', h = 0.1) {
      dS/dt = -beta*exp(w)*S*I/N
      dI/dt = beta*exp(w)*S*I/N - gamma*I
      dR/dt = gamma*I
      dZ/dt = beta*exp(w)*S*I/N
    }
  }

  sub observation {
    incidence ~ negbin(Z, k)
  }
}
