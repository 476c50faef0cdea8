// k_closed_loop.hip -- device side of the closed-loop tracking harness
// (gfx950): plant simulation, state noise, tracking error and the receding-
// horizon shift of one control step, in one single-wave kernel, so a
// control step is the batched SQP launch sequence plus this kernel and
// needs no host round trip (and can be captured in a HIP graph).
//
// Reference: track() (toplevel.cuh:505-775) with simple_simulate /
// simple_integrator_kernel (integrator.cuh:283-398) and just_shift
// (integrator.cuh:283-289).  Control step s (one knot: the plan shifts by
// one knot per step, so substeps * sim_dt is the knot interval):
//   1. the SQP (separate launches) has solved the plan xu from x_s;
//   2. plant: `substeps` explicit Euler steps of sim_dt (the reference's
//      integrator type 0: q += h qd, qd += h qdd(q, qd, u)) under the
//      plan's first control u_0 (zero-order hold);
//   3. state noise: x += sigma n, n ~ N(0, I) from splitmix64 + Box-Muller
//      (pair i of step s seeded with seed ^ 0xD1B54A32D192ED03 (8 s + i + 1));
//   4. error: ee(q_{s+1}) - ref_{s+1} (xyz), Euclidean and L1 norms;
//   5. shift: x_k <- x_{k+1}, u_k <- u_{k+1}, lambda_k <- lambda_{k+1};
//      the new tail (u_{N-2}, x_{N-1}) from the precomputed trajectory
//      (its u_{s+N-1}, x_{s+N}; the reference fills the tail from its
//      precomputed trajectory too, toplevel.cuh:680-690) or, without one,
//      u_{N-2} <- 0 and x_{N-1} kept; reference window <- path[s+1 .. s+N];
//      x_0 <- the plant state.
// Differences from the reference: the reference's plant runs the previous
// plan during the solve time (CONST_UPDATE_FREQ, SIMULATION_PERIOD) and
// shifts once per knot interval; here every control step is one knot
// interval and the new plan's u_0 is applied.  Noise: the reference's
// noise.cuh is disabled and faulty; this is the harness's own definition.
#include "common.h"
#include "model.h"

namespace mpc {

namespace {

constexpr int NT = 64;  // one wave

struct ClLds {
  float x[16];
  float sn[8], cs[8];
  float M[49];
  float bias[8];
  float qdd[8];
  float u[8];
};

__device__ __forceinline__ uint64_t sm64(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// pair i (< 7) of the step's 14 standard normals
__device__ __forceinline__ void noise_pair(uint64_t seed, uint32_t step, uint32_t i, float &a, float &b) {
  uint64_t s = seed ^ (0xD1B54A32D192ED03ull * (8ull * step + i + 1ull));
  const double u1 = (double)((sm64(s) >> 11) + 1ull) * 0x1.0p-53;  // (0, 1]
  const double u2 = (double)(sm64(s) >> 11) * 0x1.0p-53;           // [0, 1)
  const double r = sqrt(-2.0 * log(u1));
  const double t = 6.283185307179586 * u2;
  a = (float)(r * cos(t));
  b = (float)(r * sin(t));
}

__global__ void __launch_bounds__(NT) k_cl_init(uint32_t N, const float *__restrict__ x0,
                                                const float *__restrict__ path, float *__restrict__ win,
                                                float *__restrict__ xu, float *__restrict__ xp,
                                                uint32_t *__restrict__ ctr, float *__restrict__ states,
                                                float *__restrict__ rho, float rho0) {
  const uint32_t t = threadIdx.x;
  if (t < NX) {
    const float v = x0[t];
    xp[t] = v;
    xu[t] = v;
    states[t] = v;
  }
  for (uint32_t j = t; j < 6 * N; j += NT) win[j] = path[j];
  if (t == 0) {
    ctr[0] = 0;
    rho[0] = rho0;
  }
}

__global__ void __launch_bounds__(NT) k_cl_step(ClStep p, const float *__restrict__ path,
                                                const float *__restrict__ traj, float *__restrict__ win,
                                                float *__restrict__ xu, float *__restrict__ lam,
                                                float *__restrict__ xp, uint32_t *__restrict__ ctr,
                                                float *__restrict__ states, float *__restrict__ err,
                                                const uint32_t *__restrict__ pcg_now,
                                                uint32_t *__restrict__ pcg_log) {
  __shared__ ClLds L;
  const uint32_t t = threadIdx.x;
  const uint32_t step = ctr[0];
  const uint32_t N = p.N;
  if (t < NX) L.x[t] = xp[t];
  if (t < NU) L.u[t] = xu[NX + t];
  __syncthreads();
  // 2. plant: explicit Euler substeps under u_0
  for (uint32_t sub = 0; sub < p.substeps; ++sub) {
    if (t < NQ) {
      float s, c;
      sincosf(L.x[t], &s, &c);
      L.sn[t] = s;
      L.cs[t] = c;
    }
    __syncthreads();
    float s[NQ], c[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
      s[i] = L.sn[i];
      c[i] = L.cs[i];
    }
    // lanes 0..6: column t of M = RNEA(q, 0, e_t); lane 7: bias RNEA(q, qd, 0),
    // one code path for all eight lanes
    if (t <= NQ) {
      float qd[NQ], qdd[NQ], tau[NQ];
#pragma unroll
      for (int i = 0; i < NQ; ++i) {
        qd[i] = t == NQ ? L.x[NQ + i] : 0.f;
        qdd[i] = i == (int)t ? 1.f : 0.f;
      }
      rnea<false>(s, c, qd, qdd, tau, nullptr, nullptr, nullptr, nullptr, nullptr);
      if (t < NQ) {
#pragma unroll
        for (int r = 0; r < NQ; ++r) L.M[r + 7 * t] = tau[r];
      } else {
#pragma unroll
        for (int i = 0; i < NQ; ++i) L.bias[i] = tau[i];
      }
    }
    __syncthreads();
    if (t == 0) {
      float Lc[49], rhs[NQ], qdd[NQ];
      cholesky7(L.M, Lc);
#pragma unroll
      for (int i = 0; i < NQ; ++i) rhs[i] = L.u[i] - L.bias[i];
      chol_solve7(Lc, rhs, qdd);
#pragma unroll
      for (int i = 0; i < NQ; ++i) L.qdd[i] = qdd[i];
    }
    __syncthreads();
    if (t < NQ) {
      const float q = L.x[t], qd = L.x[NQ + t];
      L.x[t] = q + p.sim_dt * qd;
      L.x[NQ + t] = qd + p.sim_dt * L.qdd[t];
    }
    __syncthreads();
  }
  // 3. noise
  if (t < NQ && p.sigma > 0.f) {
    float a, b;
    noise_pair(p.seed, step, t, a, b);
    L.x[2 * t] += p.sigma * a;
    L.x[2 * t + 1] += p.sigma * b;
  }
  __syncthreads();
  if (t < NX) {
    xp[t] = L.x[t];
    states[(size_t)NX * (step + 1) + t] = L.x[t];
  }
  // 4. tracking error against the reference point of the new time
  if (t == 0) {
    float s[NQ], c[NQ], ee[3];
#pragma unroll
    for (int i = 0; i < NQ; ++i) sincosf(L.x[i], &s[i], &c[i]);
    fk_pos(s, c, ee);
    const float *r = path + (size_t)6 * (step + 1);
    const float e0 = ee[0] - r[0], e1 = ee[1] - r[1], e2 = ee[2] - r[2];
    err[2 * step] = sqrtf(e0 * e0 + e1 * e1 + e2 * e2);
    err[2 * step + 1] = fabsf(e0) + fabsf(e1) + fabsf(e2);
  }
  // 5. shift (in place, one wave: every chunk is read before it is written)
  const uint32_t nxu = NXU * N - NU, nsh = nxu - NXU;
  for (uint32_t j0 = 0; j0 < nsh; j0 += NT) {
    const uint32_t j = j0 + t;
    const float v = j < nsh ? xu[j + NXU] : 0.f;
    __syncthreads();
    if (j < nsh) xu[j] = v;
    __syncthreads();
  }
  const uint32_t nl = NX * (N - 1);
  for (uint32_t j0 = 0; j0 < nl; j0 += NT) {
    const uint32_t j = j0 + t;
    const float v = j < nl ? lam[j + NX] : 0.f;
    __syncthreads();
    if (j < nl) lam[j] = v;
    __syncthreads();
  }
  if (traj) {
    if (t < NXU) xu[NXU * (N - 2) + NX + t] = traj[(size_t)NXU * (step + N - 1) + NX + t];
  } else if (t < NU) {
    xu[NXU * (N - 2) + NX + t] = 0.f;
  }
  if (t < NX) xu[t] = L.x[t];
  for (uint32_t j = t; j < 6 * N; j += NT) win[j] = path[(size_t)6 * (step + 1) + j];
  if (pcg_log && t < p.sqp_iters) pcg_log[(size_t)p.sqp_iters * step + t] = pcg_now[t];
  __syncthreads();
  if (t == 0) ctr[0] = step + 1;
}

}  // namespace

hipError_t launch_cl_init(uint32_t N, const float *x0, const float *path, float *win, float *xu, float *xp,
                          uint32_t *ctr, float *states, float *rho, float rho0, hipStream_t st) {
  hipLaunchKernelGGL(k_cl_init, dim3(1), dim3(NT), 0, st, N, x0, path, win, xu, xp, ctr, states, rho, rho0);
  return hipGetLastError();
}

hipError_t launch_cl_step(const ClStep &p, const float *path, const float *traj, float *win, float *xu, float *lam,
                          float *xp, uint32_t *ctr, float *states, float *err, const uint32_t *pcg_now,
                          uint32_t *pcg_log, hipStream_t st) {
  if (p.N < 2 || p.sqp_iters > NT) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_cl_step, dim3(1), dim3(NT), 0, st, p, path, traj, win, xu, lam, xp, ctr, states, err,
                     pcg_now, pcg_log);
  return hipGetLastError();
}

}  // namespace mpc
