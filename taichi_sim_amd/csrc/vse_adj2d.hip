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
// and n+1 in FP32 (the differences with f32(C/d), see sumT).  Each step's
// gradient terms are formed in FP32 and added to FP32 per-cell partials; k_grad_flush
// adds the partials into the FP64 gradients after every step n divisible by
// the flush interval (32 by default), so the time accumulation is FP64 at that
// granularity while every step moves 4-byte partials instead of 8-byte totals.
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
// the adjoint differentiates exactly the arithmetic the forward ran.  No
// branch: x * 1.0f is x, and m selects the stored value at s = 1.
__device__ __forceinline__ ZCell zener_cell(const DiffArgs &A, int l, float rs) {
  ZCell z;
  const bool one = rs == 1.0f;
  z.phi_p = A.zk.phi_p[l] * rs;
  z.phi_s = A.zk.phi_s[l] * rs;
  z.h_p = A.zk.h_p[l] * rs;
  z.h_s = A.zk.h_s[l] * rs;
  z.m_p = one ? A.zk.m_p[l] : __builtin_amdgcn_rcpf(1.0f + z.h_p);
  z.m_s = one ? A.zk.m_s[l] : __builtin_amdgcn_rcpf(1.0f + z.h_s);
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

// The profile values of one axis at a band cell (profile index j), fetched
// with the cell's other operands: a wavefront holding band cells then has
// them in flight with everything else instead of waiting on them one stage
// at a time once its arithmetic has begun.  Outside the band they are unused.
struct ProfAt {
  float a, b, k, ah, bh, kh;
};
__device__ __forceinline__ ProfAt prof_at(const AxisProf &P, int j, bool band) {
  ProfAt p{0.f, 0.f, 1.f, 0.f, 0.f, 1.f};
  if (band) {
    p.a = P.a[j]; p.b = P.b[j]; p.k = P.k[j];
    p.ah = P.ah[j]; p.bh = P.bh[j]; p.kh = P.kh[j];
  }
  return p;
}

// CPML stage of a forward pass with profile values a, b, k: psi' = b psi + a D
// from psi of level n (q); returns D/k + psi' and leaves psi' in pn.
__device__ __forceinline__ float cpml_step(float a, float b, float k, float q, float D, float &pn) {
  pn = b * q + a * D;
  return D / k + pn;
}

// Adjoint of that stage: given the adjoint eb of the stretched derivative and
// qb = adjoint of psi', returns the adjoint of the raw derivative and leaves
// the adjoint of psi in qn.
__device__ __forceinline__ float cpml_back(float a, float b, float k, float qb, float eb, float &qn) {
  const float pt = qb + eb;
  qn = b * pt;
  return eb / k + a * pt;
}

__device__ __forceinline__ float ld(const VSE_G float *f, int x, int y, int nx, int ny) {
  return (x >= 0 && x < nx && y >= 0 && y < ny) ? f[(long)x * ny + y] : 0.f;
}

// Operands of one staggered difference, loaded before any arithmetic so that
// a cell's loads are all in flight together; the sums below then run the
// same operations in the same order as summing while loading.
template <int ORD>
struct Taps {
  float a[ORD], b[ORD];
};

// Along stride s: D+ has a[c] = f(i+c+1), b[c] = f(i-c); D- has a[c] = f(i+c), b[c] = f(i-c-1).
template <int ORD, bool PLUS>
__device__ __forceinline__ Taps<ORD> taps(const VSE_G float *f, long id, long s) {
  Taps<ORD> t;
#pragma unroll
  for (int c = 0; c < ORD; c++) {
    t.a[c] = PLUS ? f[id + (long)(c + 1) * s] : f[id + (long)c * s];
    t.b[c] = PLUS ? f[id - (long)c * s] : f[id - (long)(c + 1) * s];
  }
  return t;
}

// sum_c C_c (a_c - b_c) / d, the reference's derivative (wgsl2d :846-851)
template <int ORD>
__device__ __forceinline__ float diff(const DiffArgs &A, const Taps<ORD> &t, float d) {
  float acc = 0.f;
#pragma unroll
  for (int c = 0; c < ORD; c++) acc += fd_term<VSE_VAR_WGSL>(A.C[c], t.a[c], t.b[c], d, 0.f);
  return acc;
}

// Transposed staggered differences.  Forward D+ f(i) = sum C (f(i+c+1) - f(i-c)) / d
// has transpose  g(p) = sum C (h(p-c-1) - h(p+c)) / d;  D- f(i) = sum C (f(i+c) - f(i-c-1)) / d
// has transpose  g(p) = sum C (h(p-c) - h(p+c+1)) / d.  Operands outside the grid are 0.
template <int ORD, bool PLUS, int AX>
__device__ __forceinline__ Taps<ORD> tapsT(const DiffArgs &A, const VSE_G float *h, int x, int y) {
  Taps<ORD> t;
#pragma unroll
  for (int c = 0; c < ORD; c++) {
    const int o1 = PLUS ? -c - 1 : -c, o2 = PLUS ? c : c + 1;
    t.a[c] = AX == 0 ? ld(h, x + o1, y, A.nx, A.ny) : ld(h, x, y + o1, A.nx, A.ny);
    t.b[c] = AX == 0 ? ld(h, x + o2, y, A.nx, A.ny) : ld(h, x, y + o2, A.nx, A.ny);
  }
  return t;
}

// The adjoint kernels multiply by f32(C/d) instead of dividing by d: a
// rounding-level change (the gradient is checked against FP64 autograd and
// finite differences, not bit for bit), and two fewer IEEE divisions per
// operand pair.  The forward kernels keep the reference's C (a - b) / d.
template <int ORD, int AX>
__device__ __forceinline__ float sumT(const DiffArgs &A, const Taps<ORD> &t) {
  float acc = 0.f;
#pragma unroll
  for (int c = 0; c < ORD; c++) acc += (t.a[c] - t.b[c]) * A.Cd[AX][c];
  return acc;
}

// The forward's staggered difference as the adjoint recomputes it.
template <int ORD, int AX>
__device__ __forceinline__ float diffA(const DiffArgs &A, const Taps<ORD> &t) {
  float acc = 0.f;
#pragma unroll
  for (int c = 0; c < ORD; c++) acc += (t.a[c] - t.b[c]) * A.Cd[AX][c];
  return acc;
}

// Each kernel below issues every load of its cell before its first store.
// None of them reads, at another cell, anything it writes, and it reads its
// own cell's values before overwriting them, so the order of loads and stores
// changes no result; issuing the loads together keeps more of them in flight.
// LC > 0 is the mechanism count fixed at compile time, which lets the memory
// variables join the early loads; LC = 0 loops over A.L mechanisms.

// ---------------------------------------------------------------- forward
// TS = false: every relaxation-time scale is 1, so s is neither read nor
// applied (x * 1.0f == x and zener_cell selects the stored constants: the
// same bits as TS = true over a map of ones).
template <int ORD, int LC = 0, bool TS = true>
__device__ __forceinline__ void dstress_cell(const DiffArgs &A, int x, int y) {
  const int nx = A.nx, ny = A.ny;
  if (x >= nx || y >= ny) return;
  constexpr int last = ORD - 1, off = ORD - 1;
  constexpr int LR = LC > 0 ? LC : 1;
  const long id = (long)x * ny + y;
  const long ms = A.rs;          // mechanism stride of the memory-variable planes
  const float dt = A.dt;
  const DLevel &I = A.in, &O = A.out;
  const int L = A.L;
  const int ex = owned_band(A.ax[0], x, off), ey = owned_band(A.ax[1], y, off);
  const long px = (long)ex * ny + y, py = (long)x * A.ax[1].nb + ey;
  const int jx = x - off, jy = y - off;
  const bool inN = x >= last && x < nx - last - 1 && y >= last + 1 && y < ny - last;   // wgsl2d :842-912
  const bool inS = x >= last + 1 && x < nx - last && y >= last && y < ny - last - 1;   // wgsl2d :927-975

  // ---- loads
  float sxx = I.sxx[id], syy = I.syy[id], sxy = I.sxy[id];
  Taps<ORD> tvx, tvy, tyx, txy;
  float qxx = 0.f, qyy = 0.f, qyx = 0.f, qxy = 0.f;
  float rho0 = 0.f, rhoN = 0.f, rhoS = 0.f, cp0 = 0.f, cpN = 0.f, cs0 = 0.f, csN = 0.f, csS = 0.f, ts = 1.f;
  float oxx[LR], oyy[LR], oxy[LR];
  if (inN || inS) {
    rho0 = A.rho[id];
    cs0 = A.cs[id];
    if constexpr (TS) ts = A.ts[id];
  }
  if (inN) {
    tvx = taps<ORD, true>(I.vx, id, ny);
    tvy = taps<ORD, false>(I.vy, id, 1);
    if (ex >= 0) qxx = I.psi[SL_VXX][px];
    if (ey >= 0) qyy = I.psi[SL_VYY][py];
    rhoN = A.rho[id + ny];
    cpN = A.cp[id + ny];
    cp0 = A.cp[id];
    csN = A.cs[id + ny];
    if constexpr (LC > 0) {
#pragma unroll
      for (int l = 0; l < LC; l++) {
        oxx[l] = I.r[0][l * ms + id];
        oyy[l] = I.r[1][l * ms + id];
      }
    }
  }
  if (inS) {
    tyx = taps<ORD, false>(I.vy, id, ny);
    txy = taps<ORD, true>(I.vx, id, 1);
    if (ex >= 0) qyx = I.psi[SL_VYX][px];
    if (ey >= 0) qxy = I.psi[SL_VXY][py];
    rhoS = A.rho[id + 1];
    csS = A.cs[id + 1];
    if constexpr (LC > 0) {
#pragma unroll
      for (int l = 0; l < LC; l++) oxy[l] = I.r[2][l * ms + id];
    }
  }
  const ProfAt fx = prof_at(A.ax[0], jx, ex >= 0), fy = prof_at(A.ax[1], jy, ey >= 0);
  const float rs = TS ? 1.0f / ts : 1.0f;

  // ---- normal stresses
  if (inN) {
    float pn;
    float dvx = diff<ORD>(A, tvx, A.dx), dvy = diff<ORD>(A, tvy, A.dy);
    if (ex >= 0) {
      dvx = cpml_step(fx.ah, fx.bh, fx.kh, qxx, dvx, pn);
      O.psi[SL_VXX][px] = pn;
    }
    if (ey >= 0) {
      dvy = cpml_step(fy.a, fy.b, fy.k, qyy, dvy, pn);
      O.psi[SL_VYY][py] = pn;
    }
    const float rho_h = 0.5f * (rhoN + rho0);
    const float cp_h = 0.5f * (cpN + cp0);
    const float cs_l = 0.5f * (csN + cs0);
    const float cs_m = (fminf(csN, cs0) == 0.f) ? 0.f : cs_l;
    const float lambda = rho_h * (cp_h * cp_h - 2.0f * cs_l * cs_l);
    const float mu = rho_h * (cs_m * cs_m);
    const float lp2m = lambda + 2.0f * mu, lpm = lambda + mu;
    float sum_xx = 0.f, sum_yy = 0.f;
#pragma unroll
    for (int l = 0; l < (LC > 0 ? LC : L); l++) {
      const ZCell z = zener_cell(A, l, rs);
      const float ox = LC > 0 ? oxx[l] : I.r[0][l * ms + id], oy = LC > 0 ? oyy[l] : I.r[1][l * ms + id];
      const float nxx = ((ox + (dvx + dvy) * z.phi_p - ox * z.h_p) * z.m_p);
      const float nyy = ((oy + 0.5f * (dvx - dvy) * z.phi_s - oy * z.h_s) * z.m_s);
      sum_xx += ox + nxx;
      sum_yy += oy + nyy;
      O.r[0][l * ms + id] = nxx;
      O.r[1][l * ms + id] = nyy;
    }
    sxx = sxx + (lp2m * dvx + lambda * dvy + 0.5f * lpm * sum_xx + mu * sum_yy) * dt;
    syy = syy + (lambda * dvx + lp2m * dvy + 0.5f * lpm * sum_xx - mu * sum_yy) * dt;
  } else {
    for (int l = 0; l < L; l++) {
      O.r[0][l * ms + id] = I.r[0][l * ms + id];
      O.r[1][l * ms + id] = I.r[1][l * ms + id];
    }
    if (ex >= 0) O.psi[SL_VXX][px] = I.psi[SL_VXX][px];
    if (ey >= 0) O.psi[SL_VYY][py] = I.psi[SL_VYY][py];
  }

  // ---- shear stress
  if (inS) {
    float pn;
    float dvydx = diff<ORD>(A, tyx, A.dx), dvxdy = diff<ORD>(A, txy, A.dy);
    if (ex >= 0) {
      dvydx = cpml_step(fx.a, fx.b, fx.k, qyx, dvydx, pn);
      O.psi[SL_VYX][px] = pn;
    }
    if (ey >= 0) {
      dvxdy = cpml_step(fy.ah, fy.bh, fy.kh, qxy, dvxdy, pn);
      O.psi[SL_VXY][py] = pn;
    }
    const float rho_h = 0.5f * (rhoS + rho0);
    const float cs_h = (fminf(csS, cs0) == 0.f) ? 0.f : 0.5f * (csS + cs0);
    const float mu = rho_h * (cs_h * cs_h);
    float sum_xy = 0.f;
#pragma unroll
    for (int l = 0; l < (LC > 0 ? LC : L); l++) {
      const ZCell z = zener_cell(A, l, rs);
      const float o = LC > 0 ? oxy[l] : I.r[2][l * ms + id];
      const float n = ((o + (dvydx + dvxdy) * z.phi_s - o * z.h_s) * z.m_s);
      sum_xy += o + n;
      O.r[2][l * ms + id] = n;
    }
    sxy = sxy + (mu * (dvydx + dvxdy) + 0.5f * mu * sum_xy) * dt;
  } else {
    for (int l = 0; l < L; l++) O.r[2][l * ms + id] = I.r[2][l * ms + id];
    if (ex >= 0) O.psi[SL_VYX][px] = I.psi[SL_VYX][px];
    if (ey >= 0) O.psi[SL_VXY][py] = I.psi[SL_VXY][py];
  }
  O.sxx[id] = sxx;
  O.syy[id] = syy;
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
  const bool inX = x >= last + 1 && x < nx - last && y >= last + 1 && y < ny - last;   // wgsl2d :998-1025
  const bool inY = x >= last && x < nx - last - 1 && y >= last && y < ny - last - 1;   // wgsl2d :1027-1057

  // ---- loads (the source row flag first: the source-map load waits only for it)
  const int srow = A.src_rowflag[x];
  float vxn = I.vx[id], vyn = I.vy[id];
  Taps<ORD> t1, t2, t3, t4;
  float q5 = 0.f, q6 = 0.f, q7 = 0.f, q8 = 0.f;
  float rhoX = 0.f, r00 = 0.f, r10 = 0.f, r11 = 0.f, r01 = 0.f;
  if (inX) {
    t1 = taps<ORD, false>(O.sxx, id, ny);
    t2 = taps<ORD, false>(O.sxy, id, 1);
    if (ex >= 0) q5 = I.psi[SL_SXXX][px];
    if (ey >= 0) q6 = I.psi[SL_SXYY][py];
    rhoX = A.rho[id];
  }
  if (inY) {
    t3 = taps<ORD, true>(O.sxy, id, ny);
    t4 = taps<ORD, true>(O.syy, id, 1);
    if (ex >= 0) q7 = I.psi[SL_SXYX][px];
    if (ey >= 0) q8 = I.psi[SL_SYYY][py];
    r00 = A.rho[id];
    r10 = A.rho[id + ny];
    r11 = A.rho[id + ny + 1];
    r01 = A.rho[id + 1];
  }
  const ProfAt fx = prof_at(A.ax[0], x - off, ex >= 0), fy = prof_at(A.ax[1], y - off, ey >= 0);
  // source after the update, anywhere (wgsl2d :1059-1065)
  const int col = (A.n < A.src_rows && srow) ? A.src_map[id] : -1;
  const bool src_on = col >= 0;
  float rho_src = 0.f, sval = 0.f;
  if (src_on) {
    rho_src = rho_vy_edge(A, x, y, id);
    sval = A.src[(long)A.n * A.src_cols + col];
  }

  // ---- update
  if (inX) {
    float pn;
    float dx1 = diff<ORD>(A, t1, A.dx), dx2 = diff<ORD>(A, t2, A.dy);
    if (ex >= 0) {
      dx1 = cpml_step(fx.a, fx.b, fx.k, q5, dx1, pn);
      O.psi[SL_SXXX][px] = pn;
    }
    if (ey >= 0) {
      dx2 = cpml_step(fy.a, fy.b, fy.k, q6, dx2, pn);
      O.psi[SL_SXYY][py] = pn;
    }
    if (rhoX > 0.f) vxn = (dx1 + dx2) * dt / rhoX + vxn;
  } else {
    if (ex >= 0) O.psi[SL_SXXX][px] = I.psi[SL_SXXX][px];
    if (ey >= 0) O.psi[SL_SXYY][py] = I.psi[SL_SXYY][py];
  }
  if (inY) {
    float pn;
    float dy1 = diff<ORD>(A, t3, A.dx), dy2 = diff<ORD>(A, t4, A.dy);
    const float rhoY = 0.25f * (r00 + r10 + r11 + r01);
    if (ex >= 0) {
      dy1 = cpml_step(fx.ah, fx.bh, fx.kh, q7, dy1, pn);
      O.psi[SL_SXYX][px] = pn;
    }
    if (ey >= 0) {
      dy2 = cpml_step(fy.ah, fy.bh, fy.kh, q8, dy2, pn);
      O.psi[SL_SYYY][py] = pn;
    }
    if (rhoY > 0.f) vyn = (dy1 + dy2) * dt / rhoY + vyn;
  } else {
    if (ex >= 0) O.psi[SL_SXYX][px] = I.psi[SL_SXYX][px];
    if (ey >= 0) O.psi[SL_SYYY][py] = I.psi[SL_SYYY][py];
    vxn = 0.f;
    vyn = 0.f;
  }
  if (src_on && rho_src > 0.f) vyn = vyn + sval * dt / rho_src;
  O.vx[id] = vxn;
  O.vy[id] = vyn;
  return __float_as_uint(vxn * vxn + vyn * vyn);
}

// Cell of this thread in a 1D launch over the (ny/BY) x (nx/BX) tiles.  Blocks
// land on the 8 XCDs round-robin, each with its own L2; XCD k takes the
// contiguous eighth k of the tiles (when the count divides by 8), so the
// halo rows and columns of a tile mostly belong to tiles of the same XCD.
__device__ __forceinline__ void tile_cell(const DiffArgs &A, int &x, int &y) {
  const int nty = (A.ny + BY - 1) / BY;
  const int b = blockIdx.x, total = gridDim.x;
  const int t = (total & 7) == 0 ? (b & 7) * (total >> 3) + (b >> 3) : b;
  x = (t / nty) * BX + threadIdx.y;
  y = (t % nty) * BY + threadIdx.x;
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

template <int ORD, int LC, bool TS>
__global__ __launch_bounds__(BX * BY) void k_dstress(DiffArgs A) {
  int x, y;
  tile_cell(A, x, y);
  dstress_cell<ORD, LC, TS>(A, x, y);
}

template <int ORD>
__global__ __launch_bounds__(BX * BY) void k_dvelocity(DiffArgs A) {
  int x, y;
  tile_cell(A, x, y);
  const unsigned v2 = dvelocity_cell<ORD>(A, x, y);
  if (A.record && blockIdx.x == 0) record_traces(A);
  wave_max_running(v2, A.v2max);
}

// ---------------------------------------------------------------- adjoint
template <int ORD>
__global__ __launch_bounds__(BX * BY) void k_adj_velocity(DiffArgs A) {
  int x, y;
  tile_cell(A, x, y);
  const int nx = A.nx, ny = A.ny;
  // residual of step n into sigma_yy_bar at the receiver points (consumed by
  // adj_stress, which runs next; this kernel does not touch sigma_bar)
  if (blockIdx.x == 0) {
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
  const bool inXY = inX && inY;

  // ---- loads
  // adjoint of v_{n+1}: carried part + transposed stencils of the stress
  // pass of step n+1 (a1 = D+x vx, a4 = D+y vx, a2 = D-y vy, a3 = D-x vy)
  const int srow = A.src_rowflag[x];   // first: the source-map load waits only for it
  const float bvx = B.vx[id], bvy = B.vy[id];
  const Taps<ORD> u0 = tapsT<ORD, true, 0>(A, A.abar[0], x, y), u3 = tapsT<ORD, true, 1>(A, A.abar[3], x, y);
  const Taps<ORD> u1 = tapsT<ORD, false, 1>(A, A.abar[1], x, y), u2 = tapsT<ORD, false, 0>(A, A.abar[2], x, y);
  Taps<ORD> t1, t2, t3, t4;
  float p5 = 0.f, p6 = 0.f, p7 = 0.f, p8 = 0.f;       // psi' of level n+1
  float b5 = 0.f, b6 = 0.f, b7 = 0.f, b8 = 0.f;       // their adjoints
  float rhoX = 0.f, r00 = 0.f, r10 = 0.f, r11 = 0.f, r01 = 0.f, gx = 0.f, gy = 0.f;
  if (inX) {
    rhoX = A.rho[id];
    if (ex >= 0) b5 = B.psi[SL_SXXX][px];
    if (ey >= 0) b6 = B.psi[SL_SXYY][py];
  }
  if (inXY) {
    t1 = taps<ORD, false>(O.sxx, id, ny);
    t2 = taps<ORD, false>(O.sxy, id, 1);
    if (ex >= 0) p5 = O.psi[SL_SXXX][px];
    if (ey >= 0) p6 = O.psi[SL_SXYY][py];
    gx = A.gf[VSE_GR_RHO_VX][id];
  }
  if (inY) {
    r00 = A.rho[id];
    r10 = A.rho[id + ny];
    r11 = A.rho[id + ny + 1];
    r01 = A.rho[id + 1];
    t3 = taps<ORD, true>(O.sxy, id, ny);
    t4 = taps<ORD, true>(O.syy, id, 1);
    gy = A.gf[VSE_GR_RHO_VY][id];
    if (ex >= 0) {
      p7 = O.psi[SL_SXYX][px];
      b7 = B.psi[SL_SXYX][px];
    }
    if (ey >= 0) {
      p8 = O.psi[SL_SYYY][py];
      b8 = B.psi[SL_SYYY][py];
    }
  }
  const ProfAt fx = prof_at(A.ax[0], x - off, ex >= 0), fy = prof_at(A.ax[1], y - off, ey >= 0);
  const int col = (A.n < A.src_rows && srow) ? A.src_map[id] : -1;
  const bool src_on = col >= 0;
  float rho_src = 0.f, sval = 0.f;
  if (src_on) {
    rho_src = rho_vy_edge(A, x, y, id);
    sval = A.src[(long)A.n * A.src_cols + col];
  }
  if (!inY && src_on) gy = A.gf[VSE_GR_RHO_VY][id];

  // ---- adjoint of the velocity pass
  const float vxb = bvx + sumT<ORD, 0>(A, u0) + sumT<ORD, 1>(A, u3);
  const float vyb = bvy + sumT<ORD, 1>(A, u1) + sumT<ORD, 0>(A, u2);
  float a5b = 0.f, a6b = 0.f, a7b = 0.f, a8b = 0.f;
  if (inX) {
    float eb = 0.f;
    if (inY && rhoX > 0.f) {
      eb = vxb * dt / rhoX;
      const float dx1 = diffA<ORD, 0>(A, t1), dx2 = diffA<ORD, 1>(A, t2);
      // e = D/k + psi', psi' as stored in level n+1
      const float e5 = ex >= 0 ? dx1 / fx.k + p5 : dx1;
      const float e6 = ey >= 0 ? dx2 / fy.k + p6 : dx2;
      gx = gx + -(e5 + e6) * eb / rhoX;
    }
    a5b = ex >= 0 ? cpml_back(fx.a, fx.b, fx.k, b5, eb, b5) : eb;
    a6b = ey >= 0 ? cpml_back(fy.a, fy.b, fy.k, b6, eb, b6) : eb;
  }
  float ty = 0.f;
  if (inY) {
    float eb = 0.f;
    const float rhoY = 0.25f * (r00 + r10 + r11 + r01);
    if (rhoY > 0.f) {
      eb = vyb * dt / rhoY;
      const float dy1 = diffA<ORD, 0>(A, t3), dy2 = diffA<ORD, 1>(A, t4);
      const float e7 = ex >= 0 ? dy1 / fx.kh + p7 : dy1;
      const float e8 = ey >= 0 ? dy2 / fy.kh + p8 : dy2;
      ty += -(e7 + e8) * eb / rhoY;
    }
    a7b = ex >= 0 ? cpml_back(fx.ah, fx.bh, fx.kh, b7, eb, b7) : eb;
    a8b = ey >= 0 ? cpml_back(fy.ah, fy.bh, fy.kh, b8, eb, b8) : eb;
  }
  // the source term dt*src/rho_vy of step n
  if (src_on && rho_src > 0.f) ty += -(sval * dt / rho_src) * vyb / rho_src;

  // ---- stores
  if (inXY) A.gf[VSE_GR_RHO_VX][id] = gx;
  if (inY || src_on) A.gf[VSE_GR_RHO_VY][id] = gy + ty;
  if (inX) {
    if (ex >= 0) B.psi[SL_SXXX][px] = b5;
    if (ey >= 0) B.psi[SL_SXYY][py] = b6;
  }
  if (inY) {
    if (ex >= 0) B.psi[SL_SXYX][px] = b7;
    if (ey >= 0) B.psi[SL_SYYY][py] = b8;
  }
  A.abar[4][id] = a5b;
  A.abar[5][id] = a6b;
  A.abar[6][id] = a7b;
  A.abar[7][id] = a8b;
  // v_{n+1} = v_n + ... inside the vy interior, 0 outside it
  B.vx[id] = inY ? vxb : 0.f;
  B.vy[id] = inY ? vyb : 0.f;
}

template <int ORD, int LC, bool TS>
__global__ __launch_bounds__(BX * BY) void k_adj_stress(DiffArgs A) {
  int x, y;
  tile_cell(A, x, y);
  const int nx = A.nx, ny = A.ny;
  if (x >= nx || y >= ny) return;
  constexpr int last = ORD - 1, off = ORD - 1;
  constexpr int LR = LC > 0 ? LC : 1;
  const long id = (long)x * ny + y;
  const long ms = A.rs;          // mechanism stride of the memory-variable planes
  const float dt = A.dt;
  const int L = A.L;
  const DLevel &I = A.in, &B = A.bar;
  const int ex = owned_band(A.ax[0], x, off), ey = owned_band(A.ax[1], y, off);
  const long px = (long)ex * ny + y, py = (long)x * A.ax[1].nb + ey;
  const bool inN = x >= last && x < nx - last - 1 && y >= last + 1 && y < ny - last;
  const bool inS = x >= last + 1 && x < nx - last && y >= last && y < ny - last - 1;

  // ---- loads
  // adjoint of sigma_{n+1}: carried part (incl. receivers) + transposed
  // stencils of the velocity pass (a5 = D-x sxx, a6 = D-y sxy, a7 = D+x sxy, a8 = D+y syy)
  const float bsxx = B.sxx[id], bsyy = B.syy[id], bsxy = B.sxy[id];
  const Taps<ORD> u4 = tapsT<ORD, false, 0>(A, A.abar[4], x, y), u7 = tapsT<ORD, true, 1>(A, A.abar[7], x, y);
  const Taps<ORD> u5 = tapsT<ORD, false, 1>(A, A.abar[5], x, y), u6 = tapsT<ORD, true, 0>(A, A.abar[6], x, y);
  Taps<ORD> t1, t2, t3, t4;
  float q1 = 0.f, q2 = 0.f, q3 = 0.f, q4 = 0.f;       // psi of level n (v slots)
  float b1 = 0.f, b2 = 0.f, b3 = 0.f, b4 = 0.f;       // adjoints of psi'
  float rho0 = 0.f, rhoN = 0.f, rhoS = 0.f, cp0 = 0.f, cpN = 0.f, cs0 = 0.f, csN = 0.f, csS = 0.f, ts = 1.f;
  float gl = 0.f, gm = 0.f, gmxy = 0.f, gt = 0.f;
  float pxx[LR], pyy[LR], bxx[LR], byy[LR], po[LR], pb[LR];
  if (inN || inS) {
    rho0 = A.rho[id];
    cs0 = A.cs[id];
    if constexpr (TS) ts = A.ts[id];
    gt = A.gf[VSE_GR_TAU][id];
  }
  if (inN) {
    t1 = taps<ORD, true>(I.vx, id, ny);
    t2 = taps<ORD, false>(I.vy, id, 1);
    if (ex >= 0) {
      q1 = I.psi[SL_VXX][px];
      b1 = B.psi[SL_VXX][px];
    }
    if (ey >= 0) {
      q2 = I.psi[SL_VYY][py];
      b2 = B.psi[SL_VYY][py];
    }
    rhoN = A.rho[id + ny];
    cpN = A.cp[id + ny];
    cp0 = A.cp[id];
    csN = A.cs[id + ny];
    gl = A.gf[VSE_GR_LAMBDA][id];
    gm = A.gf[VSE_GR_MU][id];
    if constexpr (LC > 0) {
#pragma unroll
      for (int l = 0; l < LC; l++) {
        pxx[l] = I.r[0][l * ms + id];
        pyy[l] = I.r[1][l * ms + id];
        bxx[l] = B.r[0][l * ms + id];
        byy[l] = B.r[1][l * ms + id];
      }
    }
  }
  if (inS) {
    t3 = taps<ORD, false>(I.vy, id, ny);
    t4 = taps<ORD, true>(I.vx, id, 1);
    if (ex >= 0) {
      q3 = I.psi[SL_VYX][px];
      b3 = B.psi[SL_VYX][px];
    }
    if (ey >= 0) {
      q4 = I.psi[SL_VXY][py];
      b4 = B.psi[SL_VXY][py];
    }
    rhoS = A.rho[id + 1];
    csS = A.cs[id + 1];
    gmxy = A.gf[VSE_GR_MU_XY][id];
    if constexpr (LC > 0) {
#pragma unroll
      for (int l = 0; l < LC; l++) {
        po[l] = I.r[2][l * ms + id];
        pb[l] = B.r[2][l * ms + id];
      }
    }
  }
  const ProfAt fx = prof_at(A.ax[0], x - off, ex >= 0), fy = prof_at(A.ax[1], y - off, ey >= 0);
  const float rs = TS ? 1.0f / ts : 1.0f;

  // ---- adjoint of the stress pass
  const float P = bsxx + sumT<ORD, 0>(A, u4);
  const float Q = bsyy + sumT<ORD, 1>(A, u7);
  const float R = bsxy + sumT<ORD, 1>(A, u5) + sumT<ORD, 0>(A, u6);
  float gs = 0.f;
  float a1b = 0.f, a2b = 0.f, a3b = 0.f, a4b = 0.f;
  if (inN) {
    float pn;
    const float d1 = diffA<ORD, 0>(A, t1), d2 = diffA<ORD, 1>(A, t2);
    const float e1 = ex >= 0 ? cpml_step(fx.ah, fx.bh, fx.kh, q1, d1, pn) : d1;
    const float e2 = ey >= 0 ? cpml_step(fy.a, fy.b, fy.k, q2, d2, pn) : d2;
    const float rho_h = 0.5f * (rhoN + rho0);
    const float cp_h = 0.5f * (cpN + cp0);
    const float cs_l = 0.5f * (csN + cs0);
    const float cs_m = (fminf(csN, cs0) == 0.f) ? 0.f : cs_l;
    const float lambda = rho_h * (cp_h * cp_h - 2.0f * cs_l * cs_l);
    const float mu = rho_h * (cs_m * cs_m);
    const float lp2m = lambda + 2.0f * mu, lpm = lambda + mu;
    float e1b = dt * (lp2m * P + lambda * Q);
    float e2b = dt * (lambda * P + lp2m * Q);
    const float Sxxb = 0.5f * lpm * dt * (P + Q), Syyb = mu * dt * (P - Q);
    const float Xp = e1 + e2, Xs = 0.5f * (e1 - e2);
    float sum_xx = 0.f, sum_yy = 0.f;
#pragma unroll
    for (int l = 0; l < (LC > 0 ? LC : L); l++) {
      const ZCell z = zener_cell(A, l, rs);
      const float oxx = LC > 0 ? pxx[l] : I.r[0][l * ms + id], oyy = LC > 0 ? pyy[l] : I.r[1][l * ms + id];
      const float nxx = ((oxx + Xp * z.phi_p - oxx * z.h_p) * z.m_p);
      const float nyy = ((oyy + 0.5f * (e1 - e2) * z.phi_s - oyy * z.h_s) * z.m_s);
      sum_xx += oxx + nxx;
      sum_yy += oyy + nyy;
      const float Rxx = (LC > 0 ? bxx[l] : B.r[0][l * ms + id]) + Sxxb;      // adjoint of r'_xx
      const float Ryy = (LC > 0 ? byy[l] : B.r[1][l * ms + id]) + Syyb;
      B.r[0][l * ms + id] = Sxxb + Rxx * ((1.0f - z.h_p) * z.m_p);
      B.r[1][l * ms + id] = Syyb + Ryy * ((1.0f - z.h_s) * z.m_s);
      e1b += Rxx * z.phi_p * z.m_p + 0.5f * Ryy * z.phi_s * z.m_s;
      e2b += Rxx * z.phi_p * z.m_p - 0.5f * Ryy * z.phi_s * z.m_s;
      // d r'/d s = (m/s)(r h - X phi + r' h)
      gs += Rxx * (z.m_p * rs) * (oxx * z.h_p - Xp * z.phi_p + nxx * z.h_p);
      gs += Ryy * (z.m_s * rs) * (oyy * z.h_s - Xs * z.phi_s + nyy * z.h_s);
    }
    gl = gl + (P + Q) * (e1 + e2 + 0.5f * sum_xx) * dt;
    gm = gm + (P * (2.0f * e1 + 0.5f * sum_xx + sum_yy) + Q * (2.0f * e2 + 0.5f * sum_xx - sum_yy)) * dt;
    a1b = ex >= 0 ? cpml_back(fx.ah, fx.bh, fx.kh, b1, e1b, b1) : e1b;
    a2b = ey >= 0 ? cpml_back(fy.a, fy.b, fy.k, b2, e2b, b2) : e2b;
  }
  if (inS) {
    float pn;
    const float d3 = diffA<ORD, 0>(A, t3), d4 = diffA<ORD, 1>(A, t4);
    const float e3 = ex >= 0 ? cpml_step(fx.a, fx.b, fx.k, q3, d3, pn) : d3;
    const float e4 = ey >= 0 ? cpml_step(fy.ah, fy.bh, fy.kh, q4, d4, pn) : d4;
    const float rho_h = 0.5f * (rhoS + rho0);
    const float cs_h = (fminf(csS, cs0) == 0.f) ? 0.f : 0.5f * (csS + cs0);
    const float mu = rho_h * (cs_h * cs_h);
    const float X = e3 + e4;
    float eb = mu * dt * R;
    const float Sb = 0.5f * mu * dt * R;
    float sum_xy = 0.f;
#pragma unroll
    for (int l = 0; l < (LC > 0 ? LC : L); l++) {
      const ZCell z = zener_cell(A, l, rs);
      const float o = LC > 0 ? po[l] : I.r[2][l * ms + id];
      const float n = ((o + X * z.phi_s - o * z.h_s) * z.m_s);
      sum_xy += o + n;
      const float Rl = (LC > 0 ? pb[l] : B.r[2][l * ms + id]) + Sb;
      B.r[2][l * ms + id] = Sb + Rl * ((1.0f - z.h_s) * z.m_s);
      eb += Rl * z.phi_s * z.m_s;
      gs += Rl * (z.m_s * rs) * (o * z.h_s - X * z.phi_s + n * z.h_s);
    }
    gmxy = gmxy + R * (X + 0.5f * sum_xy) * dt;
    a3b = ex >= 0 ? cpml_back(fx.a, fx.b, fx.k, b3, eb, b3) : eb;
    a4b = ey >= 0 ? cpml_back(fy.ah, fy.bh, fy.kh, b4, eb, b4) : eb;
  }

  // ---- stores
  B.sxx[id] = P;
  B.syy[id] = Q;
  B.sxy[id] = R;
  if (inN) {
    if (ex >= 0) B.psi[SL_VXX][px] = b1;
    if (ey >= 0) B.psi[SL_VYY][py] = b2;
    A.gf[VSE_GR_LAMBDA][id] = gl;
    A.gf[VSE_GR_MU][id] = gm;
  }
  if (inS) {
    if (ex >= 0) B.psi[SL_VYX][px] = b3;
    if (ey >= 0) B.psi[SL_VXY][py] = b4;
    A.gf[VSE_GR_MU_XY][id] = gmxy;
  }
  if (inN || inS) A.gf[VSE_GR_TAU][id] = gt + gs;
  A.abar[0][id] = a1b;
  A.abar[1][id] = a2b;
  A.abar[2][id] = a3b;
  A.abar[3][id] = a4b;
}

// FP64 time accumulation: g += partial, partial = 0 (a partial that is 0
// leaves both as they are).
__global__ __launch_bounds__(256) void k_grad_flush(DiffArgs A) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i >= A.ncell) return;
#pragma unroll
  for (int k = 0; k < VSE_GR_COUNT; k++) {
    const float p = A.gf[k][i];
    if (p != 0.f) {
      A.g[k][i] += (double)p;
      A.gf[k][i] = 0.f;
    }
  }
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

dim3 grid_of(const DiffArgs &A) { return dim3((unsigned)(((A.ny + BY - 1) / BY) * ((A.nx + BX - 1) / BX))); }

}  // namespace

hipError_t vse_launch_dstress(const DiffArgs &A, hipStream_t st) {
#define CALL(O)                                                                                    \
  if (A.L == 3 && A.ts_unit) hipLaunchKernelGGL((k_dstress<O, 3, false>), grid_of(A), dim3(BY, BX), 0, st, A); \
  else if (A.L == 3) hipLaunchKernelGGL((k_dstress<O, 3, true>), grid_of(A), dim3(BY, BX), 0, st, A);      \
  else hipLaunchKernelGGL((k_dstress<O, 0, true>), grid_of(A), dim3(BY, BX), 0, st, A)
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
#define CALL(O)                                                                                       \
  if (A.L == 3 && A.ts_unit) hipLaunchKernelGGL((k_adj_stress<O, 3, false>), grid_of(A), dim3(BY, BX), 0, st, A); \
  else if (A.L == 3) hipLaunchKernelGGL((k_adj_stress<O, 3, true>), grid_of(A), dim3(BY, BX), 0, st, A);      \
  else hipLaunchKernelGGL((k_adj_stress<O, 0, true>), grid_of(A), dim3(BY, BX), 0, st, A)
  VSE_ORD_SWITCH(A.ord, CALL)
#undef CALL
  return hipGetLastError();
}

hipError_t vse_launch_grad_flush(const DiffArgs &A, hipStream_t st) {
  hipLaunchKernelGGL(k_grad_flush, dim3((unsigned)((A.ncell + 255) / 256)), dim3(256), 0, st, A);
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
