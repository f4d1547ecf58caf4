// abd_inertia.h — per-body inertia / orthogonality terms shared by the
// whole-scene and the per-scene kernels (orthogonal_energy.py).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/abd_hip.h"

namespace abd {

__device__ __forceinline__ double row_dot(const double* A, int i, int j) {
    return A[3 * i] * A[3 * j] + A[3 * i + 1] * A[3 * j + 1] + A[3 * i + 2] * A[3 * j + 2];
}
// orthogonal_energy.py:129-135
__device__ __forceinline__ void tilde(const AbdParams& P, const double* A0, const double* Adot, const double* p0,
                                      const double* pdot, double At[9], double pt[3]) {
    for (int i = 0; i < 9; ++i) At[i] = A0[i] + P.dt * Adot[i];
    const double gv[3] = {0.0, P.gravity, 0.0};
    for (int k = 0; k < 3; ++k) pt[k] = p0[k] + P.dt * pdot[k] + P.dt * P.dt * gv[k];
}
// orthogonal_energy.py:25-51, energy of body b at the state (A, p) (A[9], p[3] of body b)
__device__ __forceinline__ double inertia_energy_q(const AbdParams& P, const double* __restrict__ mass,
                                                   const double* __restrict__ I0v, int b, const double* A,
                                                   const double* p, const double* __restrict__ A0,
                                                   const double* __restrict__ p0, const double* __restrict__ Adot,
                                                   const double* __restrict__ pdot) {
    double At[9], pt[3];
    tilde(P, A0 + 9 * b, Adot + 9 * b, p0 + 3 * b, pdot + 3 * b, At, pt);
    const double m = mass ? mass[b] : P.mass, i0 = I0v ? I0v[b] : P.I0;
    double dq0[3], dq[9];
    for (int k = 0; k < 3; ++k) dq0[k] = p[k] - pt[k];
    for (int k = 0; k < 9; ++k) dq[k] = A[k] - At[k];
    const double n0 = dq0[0] * dq0[0] + dq0[1] * dq0[1] + dq0[2] * dq0[2];
    const double n1 = dq[0] * dq[0] + dq[1] * dq[1] + dq[2] * dq[2];
    const double n2 = dq[3] * dq[3] + dq[4] * dq[4] + dq[5] * dq[5];
    const double n3 = dq[6] * dq[6] + dq[7] * dq[7] + dq[8] * dq[8];
    const double dqTMdq = n0 * m + (n1 + n2 + n3) * i0;
    double eo = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double term = row_dot(A, i, j) - (i == j ? 1.0 : 0.0);
            eo += term * term;
        }
    eo = eo * P.stiffness;
    return eo * P.dt * P.dt + 0.5 * dqTMdq;
}
// the same at the stored state (Ak, pk) of body b
__device__ __forceinline__ double inertia_energy_body(const AbdParams& P, const double* __restrict__ mass,
                                                      const double* __restrict__ I0v, int b,
                                                      const double* __restrict__ Ak, const double* __restrict__ pk,
                                                      const double* __restrict__ A0, const double* __restrict__ p0,
                                                      const double* __restrict__ Adot,
                                                      const double* __restrict__ pdot) {
    return inertia_energy_q(P, mass, I0v, b, Ak + 9 * b, pk + 3 * b, A0, p0, Adot, pdot);
}
// body b's state q - alpha dq (abd.py:337-353; alpha = 0 keeps q exactly)
__device__ __forceinline__ void trial_state(int b, double alpha, const double* __restrict__ A,
                                            const double* __restrict__ p, const double* __restrict__ dq, double Ak[9],
                                            double pk[3]) {
    for (int r = 0; r < 3; ++r) pk[r] = alpha == 0.0 ? p[3 * b + r] : p[3 * b + r] - dq[3 * (4 * b) + r] * alpha;
    for (int row = 0; row < 3; ++row)
        for (int k = 0; k < 3; ++k)
            Ak[3 * row + k] = alpha == 0.0 ? A[9 * b + 3 * row + k]
                                           : A[9 * b + 3 * row + k] - dq[3 * (4 * b + 1 + row) + k] * alpha;
}

}  // namespace abd
