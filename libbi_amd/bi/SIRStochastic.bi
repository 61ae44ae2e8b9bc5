/**
 * Discrete-time stochastic epidemic: binomial infections and recoveries at
 * every unit of time, with the reported cases a binomial thinning of the
 * new infections.
 */
model SIRStochastic {
  const N = 1000.0   // population size

  param beta, gamma, rho  // infection rate, recovery rate, reporting probability
  state S, I, R, Z        // compartments; Z: new infections of the last unit of time
  noise infections, recoveries
  obs cases

  sub parameter {
    beta ~ gamma(4.0, 0.15)
    gamma ~ beta(2.0, 6.0)
    rho ~ beta(5.0, 5.0)
  }

  sub initial {
    S <- N - 10.0
    I <- 10.0
    R <- 0.0
    Z <- 0.0
  }

  sub transition(delta = 1.0) {
    infections ~ binomial(S, 1.0 - exp(-beta*I/N))
    recoveries ~ binomial(I, 1.0 - exp(-gamma))
    S <- S - infections
    I <- I + infections - recoveries
    R <- R + recoveries
    Z <- infections
  }

  sub observation {
    cases ~ binomial(Z, rho)
  }
}
