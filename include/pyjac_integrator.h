/*
 * Batched implicit integration of the species-rate ODE on the GPU.
 *
 * No counterpart in the reference: its Jacobian exists for implicit
 * integrators (the accelerInt coupling, .todo:61).  This integrates
 * d(phi)/dt = species_rates(param, phi) for every state of a batch from t = 0
 * to t_end, each state with its own adaptive step, with the two-stage
 * L-stable Rosenbrock method ROS2 (gamma = 1 + 1/sqrt(2); Verwer, Spee,
 * Blom, Hundsdorfer, SIAM J. Sci. Comput. 20 (1999) 1456):
 *
 *   (I - gamma h J) k1 = f(y_n)
 *   (I - gamma h J) k2 = f(y_n + h k1) - 2 k1
 *   y_{n+1} = y_n + 3/2 h k1 + 1/2 h k2,   error estimate h/2 (k1 + k2)
 *
 * with J the analytical Jacobian of this library, the step accepted when the
 * weighted RMS of the error estimate (weights atol + rtol max(|y_n|, |y_n+1|))
 * is at most 1, and the next step h min(5, max(0.2, 0.9 / sqrt(err))) (at
 * most h after a rejection).  States are in C order, phi[n][NN] with
 * phi = [T, V|P, n_0 .. n_{Ns-2}], param[n] = P (CONP) or V (CONV).
 */
#ifndef PYJAC_INTEGRATOR_H
#define PYJAC_INTEGRATOR_H

#include <stddef.h>

#include "pyjac_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pyjac_integrator pyjac_integrator;

typedef struct {
    long accepted;   /* accepted steps, summed over the states */
    long rejected;   /* rejected steps, summed over the states */
    long max_steps;  /* most steps (accepted + rejected) of one state */
    long failed;     /* states that did not reach t_end within max_steps */
    long launches;   /* step attempts of the batch (each state one per attempt) */
} pyjac_integrate_stats;

/* An integrator for one mechanism, NN up to 512 (PYJAC_ERR_UNSUPPORTED
   beyond).  Up to NN = 120 a state's (NN x NN) iteration matrix is factorised
   in LDS; above, in device memory, in place of its Jacobian. */
int pyjac_hip_integrator_create(const pyjac_mech_tables *tables, int conp,
                                pyjac_integrator **out);
int pyjac_hip_integrator_destroy(pyjac_integrator *it);

/*
 * Integrates n device states in place (d_phi, C order) over [0, t_end].
 * h0 > 0: the first step; h0 <= 0: 1e-2 of the step that the initial rates
 * would take to move any component by its error weight.  fixed != 0: steps
 * of exactly h0 with no error control (for convergence studies; t_end must be
 * a multiple of h0).  max_steps bounds the attempts of each state.  Blocks
 * until done; stats may be null.
 */
int pyjac_hip_integrate_dev(pyjac_integrator *it, size_t n, const double *d_param, double *d_phi,
                            double t_end, double rtol, double atol, double h0, int fixed,
                            long max_steps, void *stream, pyjac_integrate_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* PYJAC_INTEGRATOR_H */
