// k_direct.hip -- direct solve of the Schur system S lambda = gamma of every
// instance by block LDL^T (block Thomas) of the symmetric positive definite
// A = -S (gfx950).
//
// Factorisation (A_b = -S(b,b), B_b = -S(b,b-1)):
//   D_0 = A_0,   E_b = B_b D_{b-1}^-1,   D_b = A_b - E_b B_b^T,
// with D_b^-1 by Gauss-Jordan in registers (D_b is SPD: no pivoting); E_b
// and D_b^-1 are stored in the workspace (the S layout: E_b in the left
// block slot, D_b^-1 in the diagonal slot, a vector in the right slot).
// Solve of A x = r:
//   forward   y_b = r_b - E_b y_{b-1},   z_b = D_b^-1 y_b
//   backward  x_{N-1} = z_{N-1},   x_b = z_b - E_{b+1}^T x_{b+1}
// with r = -gamma, then `refine` steps of iterative refinement: the
// residual -gamma + S lambda in fp64 from the fp32 blocks, the same solve on
// it, lambda += x.  The fp32 elimination alone is off by about cond(S) x eps
// (1e-4..1e-2 relative on these systems); each step multiplies the error by
// about that factor, down to the fp32 storage of lambda.
//
// Mapping: one 16-lane group per instance, 4 groups per 64-lane workgroup
// (one wave), instances longest first.  The factorisation stages B_b and E_b
// in LDS (every group runs the wave's longest trip count; wave-level ordering
// suffices); the solve sweeps keep one vector element per lane and broadcast the
// vector inside the group by ds_swizzle, so they need no LDS and no barrier,
// and each loads the next knot's factor rows ahead.  The residual is formed
// for all knots of all instances at once, one group per knot.
//
// Reference: the QDLDL path of sqpSolve (PCG_SOLVE = 0, toplevel.cuh:160-215,
// :314-335) solves the same system by a sparse LDL^T on the host.
#include <utility>

#include "common.h"
#include "gj16.h"

namespace mpc {

namespace {

constexpr int GROUP = 16;
constexpr int GROUPS_PER_WG = 4;
constexpr int WG = GROUP * GROUPS_PER_WG;
constexpr int BPL = (NXX + GROUP - 1) / GROUP;  // B entries per lane (13)

struct FacLds {
  float B[NXX];  // B_b (column-major)
  float E[NXX];  // E_b (column-major)
};

__device__ __forceinline__ float *slot_E(float *work, size_t row) { return work + (size_t)S_STRIDE * row; }
__device__ __forceinline__ float *slot_Dinv(float *work, size_t row) { return work + (size_t)S_STRIDE * row + NXX; }
__device__ __forceinline__ float *slot_z(float *work, size_t row) { return work + (size_t)S_STRIDE * row + 2 * NXX; }

// LDS ordering among the lanes of the (single-wave) workgroup: wave_sync
// (common.h), which does not wait for outstanding global loads (the next
// knot's prefetch)

// sum_m w[m] v_m, v_m the value v of lane m of the group (m < 14), in order
template <int... M>
__device__ __forceinline__ float bdot_seq(const float (&w)[NX], float v, std::integer_sequence<int, M...>) {
  float s = 0.f;
  ((s += w[M] * bcast16<M>(v)), ...);
  return s;
}
__device__ __forceinline__ float bdot(const float (&w)[NX], float v) {
  return bdot_seq(w, v, std::make_integer_sequence<int, NX>{});
}

// instance of list slot `slot` and the workgroup's longest horizon
struct Pick {
  uint32_t inst, K0, N, Nmax;
  bool live;
};

__device__ __forceinline__ Pick pick_instance(const Batch &b, const uint32_t *order, uint32_t n_list,
                                              const uint32_t *status, uint32_t slot) {
  Pick p;
  const bool in_list = slot < n_list;
  const uint32_t s = in_list ? slot : n_list - 1;
  p.inst = order ? order[s] : s;
  p.K0 = inst_first(b, p.inst);
  p.N = inst_len(b, p.inst);
  p.live = in_list && !(status && !(status[p.inst] & ST_ACTIVE));
  const uint32_t first = (slot / GROUPS_PER_WG) * GROUPS_PER_WG;
  p.Nmax = inst_len(b, order ? order[first] : first);
  return p;
}

// factorisation sweep of one instance: E_b, D_b^-1 into the workspace
__device__ __forceinline__ void factor(FacLds &L, const float *__restrict__ S, float *__restrict__ work, const Pick &p,
                                      int lane) {
  const int j = lane < NX ? lane : NX - 1;
  float dinv[NX];  // column j of D_{b-1}^-1
#pragma unroll
  for (int r = 0; r < NX; ++r) dinv[r] = 0.f;
  float a_cur[NX], b_cur[BPL], a_nxt[NX], b_nxt[BPL];
  auto load = [&](float (&a)[NX], float (&bb)[BPL], uint32_t k) {
    const float *Sr = S + (size_t)S_STRIDE * (p.K0 + k);
#pragma unroll
    for (int r = 0; r < NX; ++r) a[r] = -Sr[NXX + r + NX * j];
#pragma unroll
    for (int t = 0; t < BPL; ++t) {
      const int i = lane + GROUP * t;
      bb[t] = (k > 0 && i < NXX) ? -Sr[i] : 0.f;
    }
  };
  load(a_cur, b_cur, 0);
  for (uint32_t k = 0; k < p.Nmax; ++k) {
    const bool on = p.live && k < p.N;
    load(a_nxt, b_nxt, k + 1 < p.N ? k + 1 : p.N - 1);
#pragma unroll
    for (int t = 0; t < BPL; ++t) {
      const int i = lane + GROUP * t;
      if (i < NXX) L.B[i] = b_cur[t];
    }
    wave_sync();
    // E_b (column j) = sum_m B(:, m) D_{b-1}^-1(m, j)
    float e[NX];
#pragma unroll
    for (int r = 0; r < NX; ++r) e[r] = 0.f;
#pragma unroll
    for (int m = 0; m < NX; ++m) {
      const float d = dinv[m];
#pragma unroll
      for (int r = 0; r < NX; ++r) e[r] += L.B[r + NX * m] * d;
    }
    if (lane < NX) {
#pragma unroll
      for (int r = 0; r < NX; ++r) L.E[r + NX * lane] = e[r];
      if (on) {
        float *w = slot_E(work, p.K0 + k) + NX * lane;
#pragma unroll
        for (int r = 0; r < NX; ++r) w[r] = e[r];
      }
    }
    wave_sync();
    // D_b (column j) = A_b(:, j) - sum_m E(:, m) B(j, m)
    float a[NX];
#pragma unroll
    for (int r = 0; r < NX; ++r) a[r] = a_cur[r];
#pragma unroll
    for (int m = 0; m < NX; ++m) {
      const float bjm = L.B[j + NX * m];
#pragma unroll
      for (int r = 0; r < NX; ++r) a[r] -= L.E[r + NX * m] * bjm;
    }
    pin(a);
    gj_inverse<NX>(lane, a, dinv);
    if (on && lane < NX) {
      float *w = slot_Dinv(work, p.K0 + k) + NX * lane;
#pragma unroll
      for (int r = 0; r < NX; ++r) w[r] = dinv[r];
    }
    wave_sync();  // every read of B and E is done before the next knot overwrites them
#pragma unroll
    for (int r = 0; r < NX; ++r) a_cur[r] = a_nxt[r];
#pragma unroll
    for (int t = 0; t < BPL; ++t) b_cur[t] = b_nxt[t];
  }
}

// ---- factorisation with one wave per instance --------------------------------
// Lane 16 q + j (q = 0..3) holds, for column j of a block, the rows
// r = 4 t + q (t = 0..3, r < 14): the products and the Gauss-Jordan updates
// of a column are split over four lanes, each element with the operation
// sequence of factor() (so E_b and D_b^-1 are bitwise the same), and the
// 1024-instance batch runs four waves per CU instead of one.  Blocks that
// every lane reads in full (B_b, E_b, D_{b-1}^-1) go through LDS.
constexpr int RS = 16;  // LDS row stride of a block (16-byte aligned rows)
struct FacWaveLds {
  float B[NX * RS];     // B_b by rows: B(r, c) at c + RS r
  float E[NX * RS];     // E_b by rows
  float Dinv[NX * RS];  // D_{b-1}^-1 by columns: (m, j) at m + RS j
};

__device__ __forceinline__ float lane_value(float v, uint32_t src) {
  return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute((int)(src << 2), __builtin_bit_cast(int, v)));
}

template <int P>
__device__ __forceinline__ void gj_wave_step(uint32_t q, uint32_t j, float (&a)[4], float (&x)[4]) {
  constexpr uint32_t QP = P & 3, TP = P >> 2;  // owner quarter and slot of row P
  float pc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) pc[t] = bcast16<P>(a[t]);  // column P, the lane's rows
  const float piv = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, a[TP]), 16 * QP + P));
  const float inv = 1.f / piv;
  const float ap = lane_value(a[TP], 16u * QP + j) * inv, xp = lane_value(x[TP], 16u * QP + j) * inv;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const uint32_t r = 4u * t + q;
    if (r == (uint32_t)P) {
      a[t] = ap;
      x[t] = xp;
    } else {
      a[t] -= pc[t] * ap;
      x[t] -= pc[t] * xp;
    }
  }
  pin(a);
  pin(x);
}

template <int P = 0>
__device__ __forceinline__ void gj_wave_steps(uint32_t q, uint32_t j, float (&a)[4], float (&x)[4]) {
  if constexpr (P < NX) {
    gj_wave_step<P>(q, j, a, x);
    gj_wave_steps<P + 1>(q, j, a, x);
  }
}

__device__ __forceinline__ void factor_wave(FacWaveLds &L, const float *__restrict__ S, float *__restrict__ work,
                                           const Pick &p, uint32_t lane) {
  const uint32_t q = lane >> 4, jl = lane & 15u, j = jl < (uint32_t)NX ? jl : NX - 1;
  const bool col_ok = jl < (uint32_t)NX;
  for (uint32_t i = lane; i < (uint32_t)(NX * RS); i += 64) L.Dinv[i] = 0.f;
  float a_cur[4], b_cur[4], a_nxt[4], b_nxt[4];
  auto load = [&](float (&a)[4], float (&bb)[4], uint32_t k) {
    const float *Sr = S + (size_t)S_STRIDE * (p.K0 + k);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const uint32_t r = 4u * t + q, i = lane + 64u * t;
      a[t] = r < (uint32_t)NX ? -Sr[NXX + r + NX * j] : 0.f;
      bb[t] = (k > 0 && i < (uint32_t)NXX) ? -Sr[i] : 0.f;
    }
  };
  load(a_cur, b_cur, 0);
  for (uint32_t k = 0; k < p.N; ++k) {
    load(a_nxt, b_nxt, k + 1 < p.N ? k + 1 : p.N - 1);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const uint32_t i = lane + 64u * t;  // B(i % 14, i / 14)
      if (i < (uint32_t)NXX) L.B[i / NX + RS * (i % NX)] = b_cur[t];
    }
    wave_sync();
    // E_b(r, j) = sum_m B(r, m) D_{b-1}^-1(m, j), m in order
    float dv[NX];
#pragma unroll
    for (int m = 0; m < NX; ++m) dv[m] = L.Dinv[m + RS * j];
    float e[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) e[t] = 0.f;
#pragma unroll
    for (int m = 0; m < NX; ++m) {
      const float d = dv[m];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const uint32_t r = 4u * t + q;
        e[t] += L.B[m + RS * (r < (uint32_t)NX ? r : 0u)] * d;
      }
    }
    if (col_ok) {
      float *w = slot_E(work, p.K0 + k) + NX * j;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const uint32_t r = 4u * t + q;
        if (r < (uint32_t)NX) {
          L.E[j + RS * r] = e[t];
          if (p.live) w[r] = e[t];
        }
      }
    }
    wave_sync();
    // D_b(r, j) = A_b(r, j) - sum_m E(r, m) B(j, m), m in order
    float a[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) a[t] = a_cur[t];
#pragma unroll
    for (int m = 0; m < NX; ++m) {
      const float bjm = L.B[m + RS * j];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const uint32_t r = 4u * t + q;
        a[t] -= L.E[m + RS * (r < (uint32_t)NX ? r : 0u)] * bjm;
      }
    }
    pin(a);
    float x[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) x[t] = 4u * t + q == j ? 1.f : 0.f;
    gj_wave_steps(q, jl, a, x);
    wave_sync();  // every read of D_{b-1}^-1, B and E is done
    if (col_ok) {
      float *w = slot_Dinv(work, p.K0 + k) + NX * j;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const uint32_t r = 4u * t + q;
        if (r < (uint32_t)NX) {
          L.Dinv[r + RS * j] = x[t];
          if (p.live) w[r] = x[t];
        }
      }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      a_cur[t] = a_nxt[t];
      b_cur[t] = b_nxt[t];
    }
  }
  wave_sync();
}

// forward and backward sweeps of one instance on the right-hand side
// r = -gamma (gamma != nullptr) or the workspace's z slot; lambda = x, or
// lambda += x with accumulate.  No barriers: the group's lanes share its
// trip count.
__device__ __forceinline__ void substitute(const float *__restrict__ gamma, float *__restrict__ work,
                                          float *__restrict__ lambda, const Pick &p, bool accumulate, int lane) {
  const int j = lane < NX ? lane : NX - 1;
  const bool wr = p.live && lane < NX;
  auto rhs = [&](uint32_t k) {
    return gamma ? -gamma[(size_t)NX * (p.K0 + k) + j] : slot_z(work, p.K0 + k)[j];
  };
  // row j of E_k (strided) and of D_k^-1 (= column j, D^-1 symmetric)
  auto load_rows = [&](float (&er)[NX], float (&dr)[NX], uint32_t k) {
    const float *E = slot_E(work, p.K0 + k), *D = slot_Dinv(work, p.K0 + k) + NX * j;
#pragma unroll
    for (int m = 0; m < NX; ++m) {
      er[m] = k > 0 ? E[j + NX * m] : 0.f;
      dr[m] = D[m];
    }
  };
  // a ring of PF buffers, knot k in buffer k % PF: each knot's rows are
  // loaded PF knots ahead and only waited for where they are used
  constexpr int PF = 4;
  float er[PF][NX], dr[PF][NX], rv[PF];
  float y = 0.f;
#pragma unroll
  for (int d = 0; d < PF; ++d)
    if ((uint32_t)d < p.N) {
      load_rows(er[d], dr[d], (uint32_t)d);
      rv[d] = rhs((uint32_t)d);
    }
  auto step = [&](const float (&e_)[NX], const float (&d_)[NX], float r, uint32_t k) {
    const float yk = r - bdot(e_, y);
    const float zk = bdot(d_, yk);
    if (wr) slot_z(work, p.K0 + k)[lane] = zk;
    y = yk;
  };
  for (uint32_t k = 0; k < p.N; k += PF) {
#pragma unroll
    for (int d = 0; d < PF; ++d) {
      if (k + d >= p.N) break;
      step(er[d], dr[d], rv[d], k + d);
      if (k + d + PF < p.N) {
        load_rows(er[d], dr[d], k + d + PF);
        rv[d] = rhs(k + d + PF);
      }
    }
  }
  // backward: column j of E_{k+1} (contiguous), through a ring the same way
  const float *zs = slot_z(work, p.K0);
  float ec[PF][NX], zv[PF];
  float x = 0.f;
  auto load_col = [&](float (&c_)[NX], float &z, int k) {  // E_{k+1} column, z_k
    const float *E = slot_E(work, p.K0 + (uint32_t)k + 1) + NX * j;
#pragma unroll
    for (int m = 0; m < NX; ++m) c_[m] = (uint32_t)k + 1 < p.N ? E[m] : 0.f;
    z = zs[(size_t)S_STRIDE * (uint32_t)k + j];
  };
#pragma unroll
  for (int d = 0; d < PF; ++d)
    if ((int)p.N - 1 - d >= 0) load_col(ec[d], zv[d], (int)p.N - 1 - d);
  auto bstep = [&](const float (&c_)[NX], float z, int k) {
    const float xk = z - bdot(c_, x);
    if (wr) {
      float *l = lambda + (size_t)NX * (p.K0 + (uint32_t)k) + lane;
      *l = accumulate ? *l + xk : xk;
    }
    x = xk;
  };
  for (int k = (int)p.N - 1; k >= 0; k -= PF) {
#pragma unroll
    for (int d = 0; d < PF; ++d) {
      if (k - d < 0) break;
      bstep(ec[d], zv[d], k - d);
      if (k - d - PF >= 0) load_col(ec[d], zv[d], k - d - PF);
    }
  }
}

__global__ void __launch_bounds__(WG) k_direct_factor_solve(Batch b, const uint32_t *__restrict__ order,
                                                            uint32_t n_list, const float *__restrict__ S,
                                                            const float *__restrict__ gamma,
                                                            const uint32_t *__restrict__ status,
                                                            float *__restrict__ work, float *__restrict__ lambda,
                                                            int do_factor, uint32_t *__restrict__ iters,
                                                            uint32_t *__restrict__ flag) {
  __shared__ FacLds lds[GROUPS_PER_WG];
  const int grp = threadIdx.x / GROUP, lane = threadIdx.x % GROUP;
  const Pick p = pick_instance(b, order, n_list, status, blockIdx.x * GROUPS_PER_WG + grp);
  if (do_factor) factor(lds[grp], S, work, p, lane);
  substitute(gamma, work, lambda, p, false, lane);
  if (p.live && lane == 0) {
    if (iters) iters[p.inst] = 0;
    if (flag) flag[p.inst] = 0;
  }
}

// one wave per instance: the factorisation on all 64 lanes, then the
// substitution sweeps on the first 16-lane group (the others follow them
// without writing)
__global__ void __launch_bounds__(64) k_direct_factor_wave(Batch b, const uint32_t *__restrict__ order,
                                                          uint32_t n_list, const float *__restrict__ S,
                                                          const float *__restrict__ gamma,
                                                          const uint32_t *__restrict__ status,
                                                          float *__restrict__ work, float *__restrict__ lambda,
                                                          uint32_t *__restrict__ iters, uint32_t *__restrict__ flag) {
  __shared__ FacWaveLds lds;
  const uint32_t lane = threadIdx.x;
  Pick p;
  p.inst = order ? order[blockIdx.x] : blockIdx.x;
  p.K0 = inst_first(b, p.inst);
  p.N = inst_len(b, p.inst);
  p.Nmax = p.N;
  p.live = !(status && !(status[p.inst] & ST_ACTIVE));
  factor_wave(lds, S, work, p, lane);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");  // the factor rows written above are read below
  Pick ps = p;
  ps.live = p.live && lane < 16u;
  substitute(gamma, work, lambda, ps, false, (int)(lane & 15u));
  if (p.live && lane == 0) {
    if (iters) iters[p.inst] = 0;
    if (flag) flag[p.inst] = 0;
  }
}

__global__ void __launch_bounds__(WG) k_direct_refine(Batch b, const uint32_t *__restrict__ order, uint32_t n_list,
                                                      const uint32_t *__restrict__ status, float *__restrict__ work,
                                                      float *__restrict__ lambda) {
  const int grp = threadIdx.x / GROUP, lane = threadIdx.x % GROUP;
  const Pick p = pick_instance(b, order, n_list, status, blockIdx.x * GROUPS_PER_WG + grp);
  substitute(nullptr, work, lambda, p, true, lane);
}

// z slot of every knot <- -gamma + S lambda (fp64 accumulation), one group per knot
__global__ void __launch_bounds__(WG) k_schur_residual(Batch b, const float *__restrict__ S,
                                                       const float *__restrict__ gamma,
                                                       const float *__restrict__ lambda, float *__restrict__ work) {
  const int grp = threadIdx.x / GROUP, lane = threadIdx.x % GROUP;
  const uint32_t gk = blockIdx.x * GROUPS_PER_WG + grp;
  if (gk >= b.total_knots || lane >= NX) return;
  const KnotRef kr = locate(b, gk);
  const float *Sr = S + (size_t)S_STRIDE * gk;
  const float *l = lambda + (size_t)NX * gk;
  double res = -(double)gamma[(size_t)NX * gk + lane];
#pragma unroll
  for (int m = 0; m < NX; ++m) res += (double)Sr[NXX + lane + NX * m] * (double)l[m];
  if (kr.k > 0) {
#pragma unroll
    for (int m = 0; m < NX; ++m) res += (double)Sr[lane + NX * m] * (double)l[m - NX];
  }
  if (kr.k + 1 < kr.N) {  // S(b,b+1) = S(b+1,b)^T: the next block row's left block
#pragma unroll
    for (int m = 0; m < NX; ++m) res += (double)Sr[S_STRIDE + m + NX * lane] * (double)l[m + NX];
  }
  slot_z(work, gk)[lane] = (float)res;
}

}  // namespace

hipError_t launch_direct_solve(const Batch &b, const uint32_t *order, uint32_t n_list, const float *S,
                               const float *gamma, const uint32_t *status, float *work, float *lambda,
                               uint32_t refine, bool do_factor, uint32_t *iters, uint32_t *flag, hipStream_t st) {
  if (n_list == 0) return hipSuccess;
  const dim3 grid((n_list + GROUPS_PER_WG - 1) / GROUPS_PER_WG);
  if (do_factor)
    hipLaunchKernelGGL(k_direct_factor_wave, dim3(n_list), dim3(64), 0, st, b, order, n_list, S, gamma, status, work,
                       lambda, iters, flag);
  else
    hipLaunchKernelGGL(k_direct_factor_solve, grid, dim3(WG), 0, st, b, order, n_list, S, gamma, status, work, lambda,
                       0, iters, flag);
  hipError_t e = hipGetLastError();
  for (uint32_t s = 0; s < refine && e == hipSuccess; ++s) {
    hipLaunchKernelGGL(k_schur_residual, dim3((b.total_knots + GROUPS_PER_WG - 1) / GROUPS_PER_WG), dim3(WG), 0, st,
                       b, S, gamma, (const float *)lambda, work);
    e = hipGetLastError();
    if (e != hipSuccess) break;
    hipLaunchKernelGGL(k_direct_refine, grid, dim3(WG), 0, st, b, order, n_list, status, work, lambda);
    e = hipGetLastError();
  }
  return e;
}

}  // namespace mpc
