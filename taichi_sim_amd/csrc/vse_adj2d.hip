// vse_adj2d.hip -- the differentiable 2D viscoelastic path.
//
// Forward: the WGSL 2D viscoelastic step (shader_webgpu_viscoelastic.wgsl
// :829-1065, the same arithmetic as vse_k2d.hip) written out of place --
// level n is read, level n+1 written -- so that the recompute sweep of the
// checkpointed gradient stores every state of a segment without copies.
// Every thread reads its own cell's state before writing it, and neighbours
// are read only from fields the pass does not write, so the kernels are also
// correct in place (in == out).  It
// adds a per-cell relaxation-time scale s: every tau of the Zener mechanisms
// at a cell is multiplied by s, so 1/tau_sigma becomes (1/tau_sigma)*(1/s)
// in phi and h (wgsl2d :881-894); at s = 1 the step is bit-identical to the
// in-place kernels.
//
// Adjoint: the exact transpose of that discrete step, applied in reverse.
// Each forward pass is  raw derivative a = D f  ->  CPML  e = a/k + psi',
// psi' = b psi + a_c a  ->  pointwise update.  Its adjoint is a pointwise
// kernel (adjoint of the update and of the CPML recursion, giving the
// adjoint of every raw derivative, abar) followed by the transposed
// stencils D^T abar, which are gathered at the start of the next adjoint
// kernel.  Per reverse step:
//   adj_velocity  v_bar += D^T(abar1..4); adjoint of the velocity pass and
//                 of the source term -> abar5..8, psi_bar (sigma slots),
//                 d/d rho_vx, rho_vy; one workgroup also adds the residual
//                 at the receiver points to sigma_yy_bar
//   adj_stress    sigma_bar += D^T(abar5..8); adjoint of the stress pass
//                 -> abar1..4, r_bar, psi_bar (v slots), d/d lambda, mu,
//                 mu_xy, tau scale
// The forward velocity kernel likewise adds the source (a per-cell map of
// source columns) and one of its workgroups records the traces, so a step
// is two launches in either direction.
// The forward values the adjoint needs (raw derivatives, CPML-stretched
// derivatives, memory variables) are recomputed from the stored levels n
// and n+1 with the forward's own FP32 arithmetic.  Gradients accumulate in
// FP64.
#include "vse_device.h"

namespace {

constexpr int BY = 64, BX = 4;
constexpr int SL_VXX = 0, SL_VYY = 1, SL_VYX = 2, SL_VXY = 3, SL_SXXX = 4, SL_SXYY = 5, SL_SXYX = 6, SL_SYYY = 7;

struct ZCell {
  float phi_p, h_p, m_p, phi_s, h_s, m_s;
};

// Zener constants of mechanism l at a cell with 1/s = rs.  At s = 1 they are
// vse_create's values (the reference's order, wgsl2d :881-894), so the step
// is bit-identical to the in-place kernels; elsewhere phi = rs * dt(1-alpha)
// / (tau_sigma sum_alpha), h = rs * dt / (2 tau_sigma), m = 1/(1+h) with a
// hardware reciprocal.  The forward and the adjoint share this function, so
// the adjoint differentiates exactly the arithmetic the forward ran.
__device__ __forceinline__ ZCell zener_cell(const DiffArgs &A, int l, float rs) {
  ZCell z;
  if (rs == 1.0f) {
    z.phi_p = A.zk.phi_p[l]; z.phi_s = A.zk.phi_s[l];
    z.h_p = A.zk.h_p[l]; z.h_s = A.zk.h_s[l];
    z.m_p = A.zk.m_p[l]; z.m_s = A.zk.m_s[l];
  } else {
    z.phi_p = A.zk.phi_p[l] * rs;
    z.phi_s = A.zk.phi_s[l] * rs;
    z.h_p = A.zk.h_p[l] * rs;
    z.h_s = A.zk.h_s[l] * rs;
    z.m_p = __builtin_amdgcn_rcpf(1.0f + z.h_p);
    z.m_s = __builtin_amdgcn_rcpf(1.0f + z.h_s);
  }
  return z;
}

// Density at the vy point with out-of-grid neighbours 0 (the source term's
// density, vse_sparse.hip; inside the vy interior it equals the update's).
__device__ __forceinline__ float rho_vy_edge(const DiffArgs &A, int x, int y, long id) {
  const bool xe = x + 1 < A.nx, ye = y + 1 < A.ny;
  return 0.25f * (A.rho[id] + (xe ? A.rho[id + A.ny] : 0.f) + ((xe && ye) ? A.rho[id + A.ny + 1] : 0.f) +
                  (ye ? A.rho[id + 1] : 0.f));
}

// Band element of grid coordinate c on an axis (profile index c - off), or -1.
__device__ __forceinline__ int owned_band(const AxisProf &P, int c, int off) {
  const int j = c - off;
  if (j < 0 || j >= P.np) return -1;
  return band_index(P, j);
}

// CPML of the out-of-place forward: psi' = b psi + a D into pout, D/k + psi'.
__device__ __forceinline__ float cpml_fwd(const VSE_G float *a, const VSE_G float *b, const VSE_G float *k,
                                          const VSE_G float *pin, VSE_G float *pout, long pidx, int j, bool band,
                                          float D) {
  if (!band) return D;
  const float pn = b[j] * pin[pidx] + a[j] * D;
  pout[pidx] = pn;
  return D / k[j] + pn;
}

// The same stretched derivative, recomputed without writing.
__device__ __forceinline__ float cpml_val(const VSE_G float *a, const VSE_G float *b, const VSE_G float *k,
                                          const VSE_G float *pin, long pidx, int j, bool band, float D) {
  if (!band) return D;
  const float pn = b[j] * pin[pidx] + a[j] * D;
  return D / k[j] + pn;
}

// Adjoint of the CPML stage: given the adjoint eb of the stretched derivative
// and psib = adjoint of psi', returns the adjoint of the raw derivative and
// leaves the adjoint of psi in psib.
__device__ __forceinline__ float cpml_adj(const VSE_G float *a, const VSE_G float *b, const VSE_G float *k,
                                          VSE_G float *psib, long pidx, int j, bool band, float eb) {
  if (!band) return eb;
  const float pt = psib[pidx] + eb;
  psib[pidx] = b[j] * pt;
  return eb / k[j] + a[j] * pt;
}

__device__ __forceinline__ float ld(const VSE_G float *f, int x, int y, int nx, int ny) {
  return (x >= 0 && x < nx && y >= 0 && y < ny) ? f[(long)x * ny + y] : 0.f;
}

// Transposed staggered differences.  Forward D+ f(i) = sum C (f(i+c+1) - f(i-c)) / d
// has transpose  g(p) = sum C (h(p-c-1) - h(p+c)) / d;  D- f(i) = sum C (f(i+c) - f(i-c-1)) / d
// has transpose  g(p) = sum C (h(p-c) - h(p+c+1)) / d.
template <int ORD, bool PLUS, int AX>
__device__ __forceinline__ float dT(const DiffArgs &A, const VSE_G float *h, int x, int y) {
  const float d = AX == 0 ? A.dx : A.dy;
  float acc = 0.f;
#pragma unroll
  for (int c = 0; c < ORD; c++) {
    const int o1 = PLUS ? -c - 1 : -c, o2 = PLUS ? c : c + 1;
    const float h1 = AX == 0 ? ld(h, x + o1, y, A.nx, A.ny) : ld(h, x, y + o1, A.nx, A.ny);
    const float h2 = AX == 0 ? ld(h, x + o2, y, A.nx, A.ny) : ld(h, x, y + o2, A.nx, A.ny);
    acc += A.C[c] * (h1 - h2) / d;
  }
  return acc;
}

// ---------------------------------------------------------------- forward
// LC > 0: compile-time mechanism count; every memory-variable load of a point
// is issued before its arithmetic (same operations and order as the runtime
// loop, so bit-identical).
template <int ORD, int LC = 0>
__device__ __forceinline__ void dstress_cell(const DiffArgs &A, int x, int y) {
  const int nx = A.nx, ny = A.ny;
  if (x >= nx || y >= ny) return;
  constexpr int last = ORD - 1, off = ORD - 1;
  const long id = (long)x * ny + y;
  const long nc = A.ncell;
  const float dt = A.dt;
  const DLevel &I = A.in, &O = A.out;
  const int L = A.L;
  const int ex = owned_band(A.ax[0], x, off), ey = owned_band(A.ax[1], y, off);
  const long px = (long)ex * ny + y, py = (long)x * A.ax[1].nb + ey;

  // ---- normal stresses (wgsl2d :842-912)
  const bool inN = x >= last && x < nx - last - 1 && y >= last + 1 && y < ny - last;
  float sxx = I.sxx[id], syy = I.syy[id];
  if (inN) {
    float dvx = 0.f, dvy = 0.f;
#pragma unroll
    for (int c = 0; c < ORD; c++) {
      dvx += fd_term<VSE_VAR_WGSL>(A.C[c], I.vx[id + (long)(c + 1) * ny], I.vx[id - (long)c * ny], A.dx, 0.f);
      dvy += fd_term<VSE_VAR_WGSL>(A.C[c], I.vy[id + c], I.vy[id - c - 1], A.dy, 0.f);
    }
    const int jx = x - off, jy = y - off;
    dvx = cpml_fwd(A.ax[0].ah, A.ax[0].bh, A.ax[0].kh, I.psi[SL_VXX], O.psi[SL_VXX], px, jx, ex >= 0, dvx);
    dvy = cpml_fwd(A.ax[1].a, A.ax[1].b, A.ax[1].k, I.psi[SL_VYY], O.psi[SL_VYY], py, jy, ey >= 0, dvy);
    const float rho_h = 0.5f * (A.rho[id + ny] + A.rho[id]);
    const float cp_h = 0.5f * (A.cp[id + ny] + A.cp[id]);
    const float csa = A.cs[id + ny], csb = A.cs[id];
    const float cs_l = 0.5f * (csa + csb);
    const float cs_m = (fminf(csa, csb) == 0.f) ? 0.f : cs_l;
    const float lambda = rho_h * (cp_h * cp_h - 2.0f * cs_l * cs_l);
    const float mu = rho_h * (cs_m * cs_m);
    const float lp2m = lambda + 2.0f * mu, lpm = lambda + mu;
    const float rs = 1.0f / A.ts[id];
    float sum_xx = 0.f, sum_yy = 0.f;
    if constexpr (LC > 0) {
      float oxx[LC], oyy[LC];
#pragma unroll
      for (int l = 0; l < LC; l++) { oxx[l] = I.r[0][l * nc + id]; oyy[l] = I.r[1][l * nc + id]; }
#pragma unroll
      for (int l = 0; l < LC; l++) {
        const ZCell z = zener_cell(A, l, rs);
        const float nxx = ((oxx[l] + (dvx + dvy) * z.phi_p - oxx[l] * z.h_p) * z.m_p);
        const float nyy = ((oyy[l] + 0.5f * (dvx - dvy) * z.phi_s - oyy[l] * z.h_s) * z.m_s);
        sum_xx += oxx[l] + nxx;
        sum_yy += oyy[l] + nyy;
        O.r[0][l * nc + id] = nxx;
        O.r[1][l * nc + id] = nyy;
      }
    } else for (int l = 0; l < L; l++) {
      const ZCell z = zener_cell(A, l, rs);
      const float oxx = I.r[0][l * nc + id], oyy = I.r[1][l * nc + id];
      const float nxx = ((oxx + (dvx + dvy) * z.phi_p - oxx * z.h_p) * z.m_p);
      const float nyy = ((oyy + 0.5f * (dvx - dvy) * z.phi_s - oyy * z.h_s) * z.m_s);
      sum_xx += oxx + nxx;
      sum_yy += oyy + nyy;
      O.r[0][l * nc + id] = nxx;
      O.r[1][l * nc + id] = nyy;
    }
    sxx = sxx + (lp2m * dvx + lambda * dvy + 0.5f * lpm * sum_xx + mu * sum_yy) * dt;
    syy = syy + (lambda * dvx + lp2m * dvy + 0.5f * lpm * sum_xx - mu * sum_yy) * dt;
  } else {
    for (int l = 0; l < L; l++) {
      O.r[0][l * nc + id] = I.r[0][l * nc + id];
      O.r[1][l * nc + id] = I.r[1][l * nc + id];
    }
    if (ex >= 0) O.psi[SL_VXX][px] = I.psi[SL_VXX][px];
    if (ey >= 0) O.psi[SL_VYY][py] = I.psi[SL_VYY][py];
  }
  O.sxx[id] = sxx;
  O.syy[id] = syy;

  // ---- shear stress (wgsl2d :927-975)
  const bool inS = x >= last + 1 && x < nx - last && y >= last && y < ny - last - 1;
  float sxy = I.sxy[id];
  if (inS) {
    float dvydx = 0.f, dvxdy = 0.f;
#pragma unroll
    for (int c = 0; c < ORD; c++) {
      dvydx += fd_term<VSE_VAR_WGSL>(A.C[c], I.vy[id + (long)c * ny], I.vy[id - (long)(c + 1) * ny], A.dx, 0.f);
      dvxdy += fd_term<VSE_VAR_WGSL>(A.C[c], I.vx[id + c + 1], I.vx[id - c], A.dy, 0.f);
    }
    const int jx = x - off, jy = y - off;
    dvydx = cpml_fwd(A.ax[0].a, A.ax[0].b, A.ax[0].k, I.psi[SL_VYX], O.psi[SL_VYX], px, jx, ex >= 0, dvydx);
    dvxdy = cpml_fwd(A.ax[1].ah, A.ax[1].bh, A.ax[1].kh, I.psi[SL_VXY], O.psi[SL_VXY], py, jy, ey >= 0, dvxdy);
    const float rho_h = 0.5f * (A.rho[id + 1] + A.rho[id]);
    const float csa = A.cs[id + 1], csb = A.cs[id];
    const float cs_h = (fminf(csa, csb) == 0.f) ? 0.f : 0.5f * (csa + csb);
    const float mu = rho_h * (cs_h * cs_h);
    const float rs = 1.0f / A.ts[id];
    float sum_xy = 0.f;
    if constexpr (LC > 0) {
      float o[LC];
#pragma unroll
      for (int l = 0; l < LC; l++) o[l] = I.r[2][l * nc + id];
#pragma unroll
      for (int l = 0; l < LC; l++) {
        const ZCell z = zener_cell(A, l, rs);
        const float n = ((o[l] + (dvydx + dvxdy) * z.phi_s - o[l] * z.h_s) * z.m_s);
        sum_xy += o[l] + n;
        O.r[2][l * nc + id] = n;
      }
    } else for (int l = 0; l < L; l++) {
      const ZCell z = zener_cell(A, l, rs);
      const float o = I.r[2][l * nc + id];
      const float n = ((o + (dvydx + dvxdy) * z.phi_s - o * z.h_s) * z.m_s);
      sum_xy += o + n;
      O.r[2][l * nc + id] = n;
    }
    sxy = sxy + (mu * (dvydx + dvxdy) + 0.5f * mu * sum_xy) * dt;
  } else {
    for (int l = 0; l < L; l++) O.r[2][l * nc + id] = I.r[2][l * nc + id];
    if (ex >= 0) O.psi[SL_VYX][px] = I.psi[SL_VYX][px];
    if (ey >= 0) O.psi[SL_VXY][py] = I.psi[SL_VXY][py];
  }
  O.sxy[id] = sxy;
}

// Velocity pass: sigma from level n+1 (just written), v and psi from level n.
template <int ORD>
__device__ __forceinline__ unsigned dvelocity_cell(const DiffArgs &A, int x, int y) {
  const int nx = A.nx, ny = A.ny;
  if (x >= nx || y >= ny) return 0u;
  constexpr int last = ORD - 1, off = ORD - 1;
  const long id = (long)x * ny + y;
  const float dt = A.dt;
  const DLevel &I = A.in, &O = A.out;
  const int ex = owned_band(A.ax[0], x, off), ey = owned_band(A.ax[1], y, off);
  const long px = (long)ex * ny + y, py = (long)x * A.ax[1].nb + ey;
  const bool inX = x >= last + 1 && x < nx - last && y >= last + 1 && y < ny - last;
  const bool inY = x >= last && x < nx - last - 1 && y >= last && y < ny - last - 1;
  float vxn = I.vx[id], vyn = I.vy[id];
  // vx (wgsl2d :998-1025)
  if (inX) {
    float d1 = 0.f, d2 = 0.f;
#pragma unroll
    for (int c = 0; c < ORD; c++) {
      d1 += fd_term<VSE_VAR_WGSL>(A.C[c], O.sxx[id + (long)c * ny], O.sxx[id - (long)(c + 1) * ny], A.dx, 0.f);
      d2 += fd_term<VSE_VAR_WGSL>(A.C[c], O.sxy[id + c], O.sxy[id - c - 1], A.dy, 0.f);
    }
    d1 = cpml_fwd(A.ax[0].a, A.ax[0].b, A.ax[0].k, I.psi[SL_SXXX], O.psi[SL_SXXX], px, x - off, ex >= 0, d1);
    d2 = cpml_fwd(A.ax[1].a, A.ax[1].b, A.ax[1].k, I.psi[SL_SXYY], O.psi[SL_SXYY], py, y - off, ey >= 0, d2);
    const float rho = A.rho[id];
    if (rho > 0.f) vxn = (d1 + d2) * dt / rho + vxn;
  } else {
    if (ex >= 0) O.psi[SL_SXXX][px] = I.psi[SL_SXXX][px];
    if (ey >= 0) O.psi[SL_SXYY][py] = I.psi[SL_SXYY][py];
  }
  // vy (wgsl2d :1027-1057)
  if (inY) {
    float d1 = 0.f, d2 = 0.f;
#pragma unroll
    for (int c = 0; c < ORD; c++) {
      d1 += fd_term<VSE_VAR_WGSL>(A.C[c], O.sxy[id + (long)(c + 1) * ny], O.sxy[id - (long)c * ny], A.dx, 0.f);
      d2 += fd_term<VSE_VAR_WGSL>(A.C[c], O.syy[id + c + 1], O.syy[id - c], A.dy, 0.f);
    }
    d1 = cpml_fwd(A.ax[0].ah, A.ax[0].bh, A.ax[0].kh, I.psi[SL_SXYX], O.psi[SL_SXYX], px, x - off, ex >= 0, d1);
    d2 = cpml_fwd(A.ax[1].ah, A.ax[1].bh, A.ax[1].kh, I.psi[SL_SYYY], O.psi[SL_SYYY], py, y - off, ey >= 0, d2);
    const float rho = 0.25f * (A.rho[id] + A.rho[id + ny] + A.rho[id + ny + 1] + A.rho[id + 1]);
    if (rho > 0.f) vyn = (d1 + d2) * dt / rho + vyn;
  } else {
    if (ex >= 0) O.psi[SL_SXYX][px] = I.psi[SL_SXYX][px];
    if (ey >= 0) O.psi[SL_SYYY][py] = I.psi[SL_SYYY][py];
    vxn = 0.f;
    vyn = 0.f;
  }
  // source after the update, anywhere (wgsl2d :1059-1065)
  const int col = A.src_map[id];
  if (col >= 0 && A.n < A.src_rows) {
    const float rho = rho_vy_edge(A, x, y, id);
    if (rho > 0.f) vyn = vyn + A.src[(long)A.n * A.src_cols + col] * dt / rho;
  }
  O.vx[id] = vxn;
  O.vy[id] = vyn;
  return __float_as_uint(vxn * vxn + vyn * vyn);
}

// Receivers of step n: sigma_yy of level n+1 summed per receiver when
// n + gate >= delay (wgsl2d :1117-1134); done by workgroup (0, 0).
__device__ __forceinline__ void record_traces(const DiffArgs &A) {
  const int t = threadIdx.x + threadIdx.y * blockDim.x;
  for (int r = t; r < A.n_rec; r += blockDim.x * blockDim.y) {
    float acc = 0.f;
    if (A.n + A.rec_gate >= A.rec_delay[r])
      for (int p = A.rec_beg[r]; p < A.rec_beg[r + 1]; p++) acc = acc + A.out.syy[A.rec_cell[p]];
    A.traces[(long)A.n * A.n_rec + r] = acc;
  }
}

template <int ORD, int LC>
__global__ __launch_bounds__(BX * BY) void k_dstress(DiffArgs A) {
  dstress_cell<ORD, LC>(A, blockIdx.y * BX + threadIdx.y, blockIdx.x * BY + threadIdx.x);
}

template <int ORD>
__global__ __launch_bounds__(BX * BY) void k_dvelocity(DiffArgs A) {
  const unsigned v2 = dvelocity_cell<ORD>(A, blockIdx.y * BX + threadIdx.y, blockIdx.x * BY + threadIdx.x);
  if (A.record && blockIdx.x == 0 && blockIdx.y == 0) record_traces(A);
  block_max_u32(v2, A.v2max);
}

// ---------------------------------------------------------------- adjoint
template <int ORD>
__global__ __launch_bounds__(BX * BY) void k_adj_velocity(DiffArgs A) {
  const int x = blockIdx.y * BX + threadIdx.y, y = blockIdx.x * BY + threadIdx.x;
  const int nx = A.nx, ny = A.ny;
  // residual of step n into sigma_yy_bar at the receiver points (consumed by
  // adj_stress, which runs next; this kernel does not touch sigma_bar)
  if (blockIdx.x == 0 && blockIdx.y == 0) {
    const int t = threadIdx.x + threadIdx.y * blockDim.x;
    for (int r = t; r < A.n_rec; r += blockDim.x * blockDim.y) {
      if (A.n + A.rec_gate < A.rec_delay[r]) continue;
      const float gb = A.gbar[(long)A.n * A.n_rec + r];
      for (int p = A.rec_beg[r]; p < A.rec_beg[r + 1]; p++) atomicAdd(&A.bar.syy[A.rec_cell[p]], gb);
    }
  }
  if (x >= nx || y >= ny) return;
  constexpr int last = ORD - 1, off = ORD - 1;
  const long id = (long)x * ny + y;
  const float dt = A.dt;
  const DLevel &O = A.out, &B = A.bar;
  const int ex = owned_band(A.ax[0], x, off), ey = owned_band(A.ax[1], y, off);
  const long px = (long)ex * ny + y, py = (long)x * A.ax[1].nb + ey;
  const bool inX = x >= last + 1 && x < nx - last && y >= last + 1 && y < ny - last;
  const bool inY = x >= last && x < nx - last - 1 && y >= last && y < ny - last - 1;

  // adjoint of v_{n+1}: carried part + transposed stencils of the stress
  // pass of step n+1 (a1 = D+x vx, a4 = D+y vx, a2 = D-y vy, a3 = D-x vy)
  const float vxb = B.vx[id] + dT<ORD, true, 0>(A, A.abar[0], x, y) + dT<ORD, true, 1>(A, A.abar[3], x, y);
  const float vyb = B.vy[id] + dT<ORD, false, 1>(A, A.abar[1], x, y) + dT<ORD, false, 0>(A, A.abar[2], x, y);

  float a5b = 0.f, a6b = 0.f, a7b = 0.f, a8b = 0.f;
  if (inX) {
    const float rho = A.rho[id];
    float eb = 0.f;
    if (inY && rho > 0.f) {
      eb = vxb * dt / rho;
      float d1 = 0.f, d2 = 0.f;
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        d1 += fd_term<VSE_VAR_WGSL>(A.C[c], O.sxx[id + (long)c * ny], O.sxx[id - (long)(c + 1) * ny], A.dx, 0.f);
        d2 += fd_term<VSE_VAR_WGSL>(A.C[c], O.sxy[id + c], O.sxy[id - c - 1], A.dy, 0.f);
      }
      // e = D/k + psi', psi' as stored in level n+1
      const float e5 = ex >= 0 ? d1 / A.ax[0].k[x - off] + O.psi[SL_SXXX][px] : d1;
      const float e6 = ey >= 0 ? d2 / A.ax[1].k[y - off] + O.psi[SL_SXYY][py] : d2;
      A.g[VSE_GR_RHO_VX][id] += -((double)(e5 + e6) * dt / ((double)rho * rho)) * vxb;
    }
    a5b = cpml_adj(A.ax[0].a, A.ax[0].b, A.ax[0].k, B.psi[SL_SXXX], px, x - off, ex >= 0, eb);
    a6b = cpml_adj(A.ax[1].a, A.ax[1].b, A.ax[1].k, B.psi[SL_SXYY], py, y - off, ey >= 0, eb);
  }
  if (inY) {
    const float rho = 0.25f * (A.rho[id] + A.rho[id + ny] + A.rho[id + ny + 1] + A.rho[id + 1]);
    float eb = 0.f;
    if (rho > 0.f) {
      eb = vyb * dt / rho;
      float d1 = 0.f, d2 = 0.f;
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        d1 += fd_term<VSE_VAR_WGSL>(A.C[c], O.sxy[id + (long)(c + 1) * ny], O.sxy[id - (long)c * ny], A.dx, 0.f);
        d2 += fd_term<VSE_VAR_WGSL>(A.C[c], O.syy[id + c + 1], O.syy[id - c], A.dy, 0.f);
      }
      const float e7 = ex >= 0 ? d1 / A.ax[0].kh[x - off] + O.psi[SL_SXYX][px] : d1;
      const float e8 = ey >= 0 ? d2 / A.ax[1].kh[y - off] + O.psi[SL_SYYY][py] : d2;
      A.g[VSE_GR_RHO_VY][id] += -((double)(e7 + e8) * dt / ((double)rho * rho)) * vyb;
    }
    a7b = cpml_adj(A.ax[0].ah, A.ax[0].bh, A.ax[0].kh, B.psi[SL_SXYX], px, x - off, ex >= 0, eb);
    a8b = cpml_adj(A.ax[1].ah, A.ax[1].bh, A.ax[1].kh, B.psi[SL_SYYY], py, y - off, ey >= 0, eb);
  }
  // the source term dt*src/rho_vy of step n
  const int col = A.src_map[id];
  if (col >= 0 && A.n < A.src_rows) {
    const float rho = rho_vy_edge(A, x, y, id);
    if (rho > 0.f) {
      const float val = A.src[(long)A.n * A.src_cols + col];
      A.g[VSE_GR_RHO_VY][id] += -((double)val * dt / ((double)rho * rho)) * vyb;
    }
  }
  A.abar[4][id] = a5b;
  A.abar[5][id] = a6b;
  A.abar[6][id] = a7b;
  A.abar[7][id] = a8b;
  // v_{n+1} = v_n + ... inside the vy interior, 0 outside it
  B.vx[id] = inY ? vxb : 0.f;
  B.vy[id] = inY ? vyb : 0.f;
}

template <int ORD, int LC>
__global__ __launch_bounds__(BX * BY) void k_adj_stress(DiffArgs A) {
  const int x = blockIdx.y * BX + threadIdx.y, y = blockIdx.x * BY + threadIdx.x;
  const int nx = A.nx, ny = A.ny;
  if (x >= nx || y >= ny) return;
  constexpr int last = ORD - 1, off = ORD - 1;
  const long id = (long)x * ny + y;
  const long nc = A.ncell;
  const float dt = A.dt;
  const int L = A.L;
  const DLevel &I = A.in, &B = A.bar;
  const int ex = owned_band(A.ax[0], x, off), ey = owned_band(A.ax[1], y, off);
  const long px = (long)ex * ny + y, py = (long)x * A.ax[1].nb + ey;

  // adjoint of sigma_{n+1}: carried part (incl. receivers) + transposed
  // stencils of the velocity pass (a5 = D-x sxx, a6 = D-y sxy, a7 = D+x sxy, a8 = D+y syy)
  const float P = B.sxx[id] + dT<ORD, false, 0>(A, A.abar[4], x, y);
  const float Q = B.syy[id] + dT<ORD, true, 1>(A, A.abar[7], x, y);
  const float R = B.sxy[id] + dT<ORD, false, 1>(A, A.abar[5], x, y) + dT<ORD, true, 0>(A, A.abar[6], x, y);
  B.sxx[id] = P;
  B.syy[id] = Q;
  B.sxy[id] = R;

  const float rs = 1.0f / A.ts[id];
  double gs = 0.0;
  float a1b = 0.f, a2b = 0.f, a3b = 0.f, a4b = 0.f;
  const bool inN = x >= last && x < nx - last - 1 && y >= last + 1 && y < ny - last;
  if (inN) {
    float d1 = 0.f, d2 = 0.f;
#pragma unroll
    for (int c = 0; c < ORD; c++) {
      d1 += fd_term<VSE_VAR_WGSL>(A.C[c], I.vx[id + (long)(c + 1) * ny], I.vx[id - (long)c * ny], A.dx, 0.f);
      d2 += fd_term<VSE_VAR_WGSL>(A.C[c], I.vy[id + c], I.vy[id - c - 1], A.dy, 0.f);
    }
    const float e1 = cpml_val(A.ax[0].ah, A.ax[0].bh, A.ax[0].kh, I.psi[SL_VXX], px, x - off, ex >= 0, d1);
    const float e2 = cpml_val(A.ax[1].a, A.ax[1].b, A.ax[1].k, I.psi[SL_VYY], py, y - off, ey >= 0, d2);
    const float rho_h = 0.5f * (A.rho[id + ny] + A.rho[id]);
    const float cp_h = 0.5f * (A.cp[id + ny] + A.cp[id]);
    const float csa = A.cs[id + ny], csb = A.cs[id];
    const float cs_l = 0.5f * (csa + csb);
    const float cs_m = (fminf(csa, csb) == 0.f) ? 0.f : cs_l;
    const float lambda = rho_h * (cp_h * cp_h - 2.0f * cs_l * cs_l);
    const float mu = rho_h * (cs_m * cs_m);
    const float lp2m = lambda + 2.0f * mu, lpm = lambda + mu;
    float e1b = dt * (lp2m * P + lambda * Q);
    float e2b = dt * (lambda * P + lp2m * Q);
    const float Sxxb = 0.5f * lpm * dt * (P + Q), Syyb = mu * dt * (P - Q);
    const float Xp = e1 + e2, Xs = 0.5f * (e1 - e2);
    float sum_xx = 0.f, sum_yy = 0.f;
    float pxx[LC > 0 ? LC : 1], pyy[LC > 0 ? LC : 1], bxx[LC > 0 ? LC : 1], byy[LC > 0 ? LC : 1];
    if constexpr (LC > 0) {
#pragma unroll
      for (int l = 0; l < LC; l++) {
        pxx[l] = I.r[0][l * nc + id]; pyy[l] = I.r[1][l * nc + id];
        bxx[l] = B.r[0][l * nc + id]; byy[l] = B.r[1][l * nc + id];
      }
    }
#pragma unroll
    for (int l = 0; l < (LC > 0 ? LC : L); l++) {
      const ZCell z = zener_cell(A, l, rs);
      const float oxx = LC > 0 ? pxx[l] : I.r[0][l * nc + id], oyy = LC > 0 ? pyy[l] : I.r[1][l * nc + id];
      const float nxx = ((oxx + Xp * z.phi_p - oxx * z.h_p) * z.m_p);
      const float nyy = ((oyy + 0.5f * (e1 - e2) * z.phi_s - oyy * z.h_s) * z.m_s);
      sum_xx += oxx + nxx;
      sum_yy += oyy + nyy;
      const float Rxx = (LC > 0 ? bxx[l] : B.r[0][l * nc + id]) + Sxxb;      // adjoint of r'_xx
      const float Ryy = (LC > 0 ? byy[l] : B.r[1][l * nc + id]) + Syyb;
      B.r[0][l * nc + id] = Sxxb + Rxx * ((1.0f - z.h_p) * z.m_p);
      B.r[1][l * nc + id] = Syyb + Ryy * ((1.0f - z.h_s) * z.m_s);
      e1b += Rxx * z.phi_p * z.m_p + 0.5f * Ryy * z.phi_s * z.m_s;
      e2b += Rxx * z.phi_p * z.m_p - 0.5f * Ryy * z.phi_s * z.m_s;
      // d r'/d s = (m/s)(r h - X phi + r' h)
      gs += (double)Rxx * ((double)z.m_p * rs) * ((double)oxx * z.h_p - (double)Xp * z.phi_p + (double)nxx * z.h_p);
      gs += (double)Ryy * ((double)z.m_s * rs) * ((double)oyy * z.h_s - (double)Xs * z.phi_s + (double)nyy * z.h_s);
    }
    A.g[VSE_GR_LAMBDA][id] += ((double)P + Q) * ((double)e1 + e2 + 0.5 * sum_xx) * dt;
    A.g[VSE_GR_MU][id] += ((double)P * (2.0 * e1 + 0.5 * sum_xx + sum_yy) +
                           (double)Q * (2.0 * e2 + 0.5 * sum_xx - sum_yy)) * dt;
    a1b = cpml_adj(A.ax[0].ah, A.ax[0].bh, A.ax[0].kh, B.psi[SL_VXX], px, x - off, ex >= 0, e1b);
    a2b = cpml_adj(A.ax[1].a, A.ax[1].b, A.ax[1].k, B.psi[SL_VYY], py, y - off, ey >= 0, e2b);
  }
  const bool inS = x >= last + 1 && x < nx - last && y >= last && y < ny - last - 1;
  if (inS) {
    float d3 = 0.f, d4 = 0.f;
#pragma unroll
    for (int c = 0; c < ORD; c++) {
      d3 += fd_term<VSE_VAR_WGSL>(A.C[c], I.vy[id + (long)c * ny], I.vy[id - (long)(c + 1) * ny], A.dx, 0.f);
      d4 += fd_term<VSE_VAR_WGSL>(A.C[c], I.vx[id + c + 1], I.vx[id - c], A.dy, 0.f);
    }
    const float e3 = cpml_val(A.ax[0].a, A.ax[0].b, A.ax[0].k, I.psi[SL_VYX], px, x - off, ex >= 0, d3);
    const float e4 = cpml_val(A.ax[1].ah, A.ax[1].bh, A.ax[1].kh, I.psi[SL_VXY], py, y - off, ey >= 0, d4);
    const float rho_h = 0.5f * (A.rho[id + 1] + A.rho[id]);
    const float csa = A.cs[id + 1], csb = A.cs[id];
    const float cs_h = (fminf(csa, csb) == 0.f) ? 0.f : 0.5f * (csa + csb);
    const float mu = rho_h * (cs_h * cs_h);
    const float X = e3 + e4;
    float eb = mu * dt * R;
    const float Sb = 0.5f * mu * dt * R;
    float sum_xy = 0.f;
    float po[LC > 0 ? LC : 1], pb[LC > 0 ? LC : 1];
    if constexpr (LC > 0) {
#pragma unroll
      for (int l = 0; l < LC; l++) { po[l] = I.r[2][l * nc + id]; pb[l] = B.r[2][l * nc + id]; }
    }
#pragma unroll
    for (int l = 0; l < (LC > 0 ? LC : L); l++) {
      const ZCell z = zener_cell(A, l, rs);
      const float o = LC > 0 ? po[l] : I.r[2][l * nc + id];
      const float n = ((o + X * z.phi_s - o * z.h_s) * z.m_s);
      sum_xy += o + n;
      const float Rl = (LC > 0 ? pb[l] : B.r[2][l * nc + id]) + Sb;
      B.r[2][l * nc + id] = Sb + Rl * ((1.0f - z.h_s) * z.m_s);
      eb += Rl * z.phi_s * z.m_s;
      gs += (double)Rl * ((double)z.m_s * rs) * ((double)o * z.h_s - (double)X * z.phi_s + (double)n * z.h_s);
    }
    A.g[VSE_GR_MU_XY][id] += (double)R * ((double)X + 0.5 * sum_xy) * dt;
    a3b = cpml_adj(A.ax[0].a, A.ax[0].b, A.ax[0].k, B.psi[SL_VYX], px, x - off, ex >= 0, eb);
    a4b = cpml_adj(A.ax[1].ah, A.ax[1].bh, A.ax[1].kh, B.psi[SL_VXY], py, y - off, ey >= 0, eb);
  }
  if (inN || inS) A.g[VSE_GR_TAU][id] += gs;
  A.abar[0][id] = a1b;
  A.abar[1][id] = a2b;
  A.abar[2][id] = a3b;
  A.abar[3][id] = a4b;
}

__global__ __launch_bounds__(256) void k_src_map(int *map, long ncell, const int *cells, const int *cols, int n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < ncell) map[i] = -1;
}

__global__ __launch_bounds__(256) void k_src_map_set(int *map, const int *cells, const int *cols, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) map[cells[i]] = cols[i];
}

// 0.5 * sum (traces - obs)^2 in FP64; resid = traces - obs
__global__ __launch_bounds__(256) void k_misfit(const float *tr, const float *obs, float *resid, long n, double *acc) {
  double s = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float d = tr[i] - obs[i];
    resid[i] = d;
    s += 0.5 * (double)d * (double)d;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  __shared__ double ws[4];
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(acc, ws[0] + ws[1] + ws[2] + ws[3]);
}

#define VSE_ORD_SWITCH(ord, CALL)                                                                      \
  switch (ord) {                                                                                       \
    case 2: CALL(2); break;                                                                            \
    case 3: CALL(3); break;                                                                            \
    case 4: CALL(4); break;                                                                            \
    case 5: CALL(5); break;                                                                            \
    case 6: CALL(6); break;                                                                            \
    default: return hipErrorInvalidValue;                                                              \
  }

dim3 grid_of(const DiffArgs &A) { return dim3((A.ny + BY - 1) / BY, (A.nx + BX - 1) / BX); }

}  // namespace

hipError_t vse_launch_dstress(const DiffArgs &A, hipStream_t st) {
#define CALL(O)                                                                                 \
  if (A.L == 3) hipLaunchKernelGGL((k_dstress<O, 3>), grid_of(A), dim3(BY, BX), 0, st, A);      \
  else hipLaunchKernelGGL((k_dstress<O, 0>), grid_of(A), dim3(BY, BX), 0, st, A)
  VSE_ORD_SWITCH(A.ord, CALL)
#undef CALL
  return hipGetLastError();
}

hipError_t vse_launch_dvelocity(const DiffArgs &A, hipStream_t st) {
#define CALL(O) hipLaunchKernelGGL((k_dvelocity<O>), grid_of(A), dim3(BY, BX), 0, st, A)
  VSE_ORD_SWITCH(A.ord, CALL)
#undef CALL
  return hipGetLastError();
}

hipError_t vse_launch_adj_velocity(const DiffArgs &A, hipStream_t st) {
#define CALL(O) hipLaunchKernelGGL((k_adj_velocity<O>), grid_of(A), dim3(BY, BX), 0, st, A)
  VSE_ORD_SWITCH(A.ord, CALL)
#undef CALL
  return hipGetLastError();
}

hipError_t vse_launch_adj_stress(const DiffArgs &A, hipStream_t st) {
#define CALL(O)                                                                                 \
  if (A.L == 3) hipLaunchKernelGGL((k_adj_stress<O, 3>), grid_of(A), dim3(BY, BX), 0, st, A);   \
  else hipLaunchKernelGGL((k_adj_stress<O, 0>), grid_of(A), dim3(BY, BX), 0, st, A)
  VSE_ORD_SWITCH(A.ord, CALL)
#undef CALL
  return hipGetLastError();
}

hipError_t vse_launch_src_map(int *map, long ncell, const int *cells, const int *cols, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_src_map, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, st, map, ncell, cells, cols, n);
  if (n > 0) hipLaunchKernelGGL(k_src_map_set, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, map, cells, cols, n);
  return hipGetLastError();
}

hipError_t vse_launch_misfit(const float *traces, const float *obs, float *resid, long n, double *acc,
                             hipStream_t st) {
  long blocks = (n + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_misfit, dim3((unsigned)blocks), dim3(256), 0, st, traces, obs, resid, n, acc);
  return hipGetLastError();
}
