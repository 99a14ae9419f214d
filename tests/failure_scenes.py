"""Scenes that raise each of the reference's run-stopping errors on their first step (shared by the oracle and
engine failure tests). Each is a 216-particle block (parity_scenes.model_block) whose parameters drive the
return map into the error branch:
  * Perzyna increment below zero (plasticity.cpp:190-193): a negative yield stress makes the closed-form
    denominator f_mu + q_y eta / dt negative;
  * Newton iteration limit (:200-203): a viscosity so large that the rate term stalls the residual above 1e-2;
  * Newton derivative below 1e-14 (:219-222): a larger viscosity still, where (dt / (eta dgamma + dt))^n vanishes;
  * MCC without an admissible step (mcc_rma_implicit_sinh_onevar.hpp): a steep sinh hardening where the limited
    line search fails both ways (parameters found by a seeded search over the oracle)."""
import parity_scenes as PS

FAILURES = {
    "perzyna_negative": (lambda: PS.model_block("VMVisc", n=6, q_max=-1e3, q_min=-1e3, perzyna_visc=1.0, xi=0.0),
                         "PERZYNA_NEGATIVE"),
    "newton_maxiter": (lambda: PS.model_block("VMVisc", n=6, q_max=400.0, q_min=400.0, perzyna_visc=1e9, perzyna_exp=2.0,
                                              xi=0.5), "NEWTON_MAXITER"),
    "newton_small_deriv": (lambda: PS.model_block("VMVisc", n=6, q_max=400.0, q_min=400.0, perzyna_visc=1e12,
                                                  perzyna_exp=2.0, xi=0.5), "NEWTON_SMALL_DERIV"),
    "dpvisc_maxiter": (lambda: PS.model_block("DPVisc", n=6, perzyna_visc=1e9, perzyna_exp=2.0), "NEWTON_MAXITER"),
    "mcc_no_step": (lambda: PS.model_block("MCC", n=6, hardening="SinhImpl", M=1.42, p0=6229.0, beta=0.19, xi=16.0),
                    "MCC_NO_STEP"),
}


def build(lib, name):
    import mpm_abi as A
    cfg, plates, objs, f = FAILURES[name][0]()
    e = A.Engine(lib, cfg, plates, objs)
    e.set_particles(f["x"], **{k: v for k, v in f.items() if k != "x"})
    return e
