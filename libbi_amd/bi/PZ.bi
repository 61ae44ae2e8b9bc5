/**
 * Predator-prey plankton model: the equations, constants and distributions
 * of the LibBi user guide's first example (docs/src/users.tex:52-84), with
 * an explicit integrator (RK4, step 0.1) in the ode block, since the guide's
 * default RK4(3) is not built here.
 *
 * P grows at a daily random rate alpha and is grazed by Z; Z converts a
 * fraction e of what it grazes into growth and dies at linear and quadratic
 * rates.  Only log P is observed, with noise of standard deviation 0.2.
 */
model PZ {
  const c = 0.25    // grazing (clearance) rate of Z on P
  const e = 0.3     // fraction of grazed P that becomes Z
  const m_l = 0.1   // linear death rate of Z
  const m_q = 0.1   // quadratic death rate of Z

  param mu, sigma   // daily growth rate of P: its mean and its spread
  state P, Z        // prey (phytoplankton) and predator (zooplankton)
  noise alpha       // the day's growth rate of P
  obs P_obs         // measured P

  sub parameter {
    mu ~ uniform(0.0, 1.0)
    sigma ~ uniform(0.0, 0.5)
  }

  sub initial {
    P ~ log_normal(log(2.0), 0.2)
    Z ~ log_normal(log(2.0), 0.1)
  }

  sub transition {
    alpha ~ normal(mu, sigma)
    ode(alg = 'RK4', h = 0.1) {
      dP/dt = alpha*P - c*P*Z
      dZ/dt = e*c*P*Z - m_l*Z - m_q*Z*Z
    }
  }

  sub observation {
    P_obs ~ log_normal(log(P), 0.2)
  }
}
