// vse_k3d.hip -- 3D stress and velocity kernels (elastic / Zener, CPML).
//
// Update semantics follow shader_webgpu_3D_viscoelastic.wgsl:1340-1773; the
// CPU oracle restates the same order and every expression keeps it.
//
// Memory plan (both passes are HBM-bound stencils, ~1.4 flop/byte):
// * a workgroup owns a (TY x 64) column tile in (y, z) -- z is the contiguous
//   axis, so each wavefront reads 256 B per field -- and marches along x (the
//   slowest axis, plane stride ny*nz) over a chunk of planes;
// * the x-direction stencil window of each thread's own column lives in
//   registers and advances one plane per iteration, so every plane of a
//   stencil field is fetched once instead of 2*ORD times;
// * y/z neighbours at the current plane are plain loads that hit L1/L2: the
//   tile itself fetched them this iteration, and tiles adjacent in y run on
//   the same XCD (tile index remapped so each XCD owns contiguous tiles);
// * fields touched only at the thread's own cell (memory variables r, the
//   updated field, psi, cp) are streamed with non-temporal loads/stores so
//   they do not evict the stencil planes from L2.
#include <algorithm>

#include "vse_device.h"

namespace {

constexpr int BZ = 64;

// One x-plane of a field as a raw buffer: a scalar (SGPR) resource
// descriptor over the plane's bytes, addressed with a 32-bit per-lane offset.
// The many loads in flight then cost one VGPR of address each (flat 64-bit
// addresses cost two plus the add chains), and the hardware range-checks
// every access against the plane (out-of-range reads return 0, writes drop).
// Word 3 0x00020000 is the gfx9-family raw 32-bit buffer format.  CP = 2 sets
// the non-temporal cache policy for pure streams.
struct PB {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ PB(const float *plane_base, unsigned nbytes)
      : r(__builtin_amdgcn_make_buffer_rsrc((void *)plane_base, (short)0, (int)nbytes, 0x00020000)) {}
  template <int CP = 0>
  __device__ __forceinline__ float ld(unsigned i) const {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)(i * 4u), 0, CP));
  }
  template <int CP = 0>
  __device__ __forceinline__ void st(unsigned i, float v) const {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)(i * 4u), 0, CP);
  }
};

// A whole field (fewer than 2^30 cells) as one raw buffer, indexed by
// plane * sx + in-plane offset in elements.  The descriptor is built once per
// kernel, so a plane step costs vector offset adds shared by every field
// instead of a scalar descriptor per plane and field.  Every access is
// range-checked against the field: an index past its end, or a negative one
// (which wraps above it), reads 0.
struct FBuf {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ FBuf(const float *base, long ncell)
      : r(__builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)(unsigned)(ncell * 4), 0x00020000)) {}
  template <int CP = 0>
  __device__ __forceinline__ float ld(unsigned i) const {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)(i * 4u), 0, CP));
  }
  template <int CP = 0>
  __device__ __forceinline__ void st(unsigned i, float v) const {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)(i * 4u), 0, CP);
  }
};

// psi recursion psi <- b psi + a D, returns D/k + psi (wgsl3d :1372-1378);
// FAST multiplies by the precomputed 1/k.
template <int CP, bool FAST>
__device__ __forceinline__ float stretch(const float *a, const float *b, const float *k, const float *ki,
                                         const PB &P, unsigned i, int j, float D) {
  const float pn = b[j] * P.ld<CP>(i) + a[j] * D;
  P.st<CP>(i, pn);
  if constexpr (FAST) return D * ki[j] + pn;
  else return D / k[j] + pn;
}

// One staggered-difference term along axis AX: C*(a-b)/d (reference order)
// or (a-b)*f32(C/d) (fast).
template <bool FAST, int AX>
__device__ __forceinline__ float dterm(const StepArgs &A, int c, float a, float b) {
  if constexpr (FAST) return (a - b) * A.Cd[AX][c];
  else return A.C[c] * (a - b) / (AX == 0 ? A.dx : (AX == 1 ? A.dy : A.dz));
}

// LC > 0: the number of Zener mechanisms is the compile-time LC, and every
// own-cell load of a section (memory variables, then stresses) is issued
// before its arithmetic, so a wavefront has them all in flight at once
// (the runtime-L loop waits on each mechanism's loads in turn).  The
// arithmetic and its order are the same, so results are bit-identical.
template <bool VISCO, int ORD, int TY, bool NT, bool FAST, int LC = 0>
__global__ __launch_bounds__(BZ * TY) void k_stress3d(StepArgs A, int xchunk, int nzb, int nyb, int nxc, int remap) {
  constexpr int CP = NT ? 2 : 0;
  const Tile T = tile_of_block(nzb, nyb, nxc, remap);
  const int z = T.bz * BZ + threadIdx.x;
  const int y = T.by * TY + threadIdx.y;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  if (z >= nz || y >= ny) return;
  constexpr int last = ORD - 1, off = ORD - 1;
  const unsigned sy = nz;
  const long sx = (long)ny * nz;
  const unsigned pbytes = (unsigned)(sx * 4);
  const unsigned col = y * sy + z;
  const float dt = A.dt;
  const int x0 = T.bxc * xchunk, x1 = min(nx, x0 + xchunk);
  const int jy = y - off, jz = z - off;
  const int by = band_index(A.ax[1], jy), bz = band_index(A.ax[2], jz);
  const int nby = A.ax[1].nb, nbz = A.ax[2].nb;
  const unsigned py_col = by * nz + z;      // offset in a psi_y plane (nby * nz per x)
  const unsigned pz_col = y * nbz + bz;     // offset in a psi_z plane (ny * nbz per x)
  const unsigned pyb = (unsigned)nby * nz * 4u, pzb = (unsigned)ny * nbz * 4u;
  const bool yp = y >= last && y < ny - last - 1;
  const bool zp = z >= last && z < nz - last - 1;
  const bool ym = y >= last + 1 && y < ny - last;
  const bool zm = z >= last + 1 && z < nz - last;
  auto P = [&](const float *f, int x) { return PB(f + (long)(x < 0 ? 0 : (x >= nx ? nx - 1 : x)) * sx, pbytes); };

  // x windows of the own column: vx at x-(ORD-1)..x+ORD, vy/vz at x-ORD..x+ORD-1
  // (planes outside [0,nx) read clamped values that no update uses)
  float wvx[2 * ORD], wvy[2 * ORD], wvz[2 * ORD];
#pragma unroll
  for (int i = 0; i < 2 * ORD - 1; i++) {
    wvx[i + 1] = P(A.vx, x0 - (ORD - 1) + i).ld(col);
    wvy[i + 1] = P(A.vy, x0 - ORD + i).ld(col);
    wvz[i + 1] = P(A.vz, x0 - ORD + i).ld(col);
  }
  float rho_n = P(A.rho, x0).ld(col), cs_n = P(A.cs, x0).ld(col);

#pragma unroll 1
  for (int x = x0; x < x1; x++) {
#pragma unroll
    for (int i = 0; i < 2 * ORD - 1; i++) { wvx[i] = wvx[i + 1]; wvy[i] = wvy[i + 1]; wvz[i] = wvz[i + 1]; }
    wvx[2 * ORD - 1] = P(A.vx, x + ORD).ld(col);
    wvy[2 * ORD - 1] = P(A.vy, x + ORD - 1).ld(col);
    wvz[2 * ORD - 1] = P(A.vz, x + ORD - 1).ld(col);
    const float rho0 = rho_n, cs0 = cs_n;
    const PB rho1P = P(A.rho, x + 1), cs1P = P(A.cs, x + 1);
    rho_n = rho1P.ld(col);
    cs_n = cs1P.ld(col);
    const int jx = x - off;
    const int bx = band_index(A.ax[0], jx);
    const bool xp = x >= last && x < nx - last - 1;
    const bool xm = x >= last + 1 && x < nx - last;
    const PB vxP = P(A.vx, x), vyP = P(A.vy, x), vzP = P(A.vz, x), rhoP = P(A.rho, x), csP = P(A.cs, x);
    const long rofs = (long)x * sx;
    // LC > 0: the plane's neighbour loads (v along y and z, rho/cs next to
    // the shear points) are all issued here, before any arithmetic; buffer
    // loads outside a plane read 0 and those values are used only outside
    // the interiors, where they are not.
    constexpr bool PRE = VISCO && LC > 0;
    float nvy_p[ORD], nvy_m[ORD], nvz_p[ORD], nvz_m[ORD];        // normal: dvy (D-), dvz (D-)
    float nxy_p[ORD], nxy_m[ORD], nxz_p[ORD], nxz_m[ORD];        // shear: dvx/dy, dvx/dz (D+)
    float nzy_p[ORD], nzy_m[ORD], nyz_p[ORD], nyz_m[ORD];        // shear: dvz/dy, dvy/dz (D+)
    float m_r0y = 0.f, m_r1y = 0.f, m_c0y = 0.f, m_c1y = 0.f, m_r0z = 0.f, m_r1z = 0.f, m_c0z = 0.f, m_c1z = 0.f;
    float m_r0yz = 0.f, m_c0yz = 0.f, m_cp = 0.f;
    if constexpr (PRE) {
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        // the own cell of plane x is already in the x windows
        nvy_p[c] = c == 0 ? wvy[ORD] : vyP.ld(col + c * sy); nvy_m[c] = vyP.ld(col - (c + 1) * sy);
        nvz_p[c] = c == 0 ? wvz[ORD] : vzP.ld(col + c); nvz_m[c] = vzP.ld(col - c - 1);
        nxy_p[c] = vxP.ld(col + (c + 1) * sy); nxy_m[c] = c == 0 ? wvx[ORD - 1] : vxP.ld(col - c * sy);
        nxz_p[c] = vxP.ld(col + c + 1); nxz_m[c] = c == 0 ? wvx[ORD - 1] : vxP.ld(col - c);
        nzy_p[c] = vzP.ld(col + (c + 1) * sy); nzy_m[c] = c == 0 ? wvz[ORD] : vzP.ld(col - c * sy);
        nyz_p[c] = vyP.ld(col + c + 1); nyz_m[c] = c == 0 ? wvy[ORD] : vyP.ld(col - c);
      }
      m_r0y = rhoP.ld(col + sy); m_r1y = rho1P.ld(col + sy); m_c0y = csP.ld(col + sy); m_c1y = cs1P.ld(col + sy);
      m_r0z = rhoP.ld(col + 1); m_r1z = rho1P.ld(col + 1); m_c0z = csP.ld(col + 1); m_c1z = cs1P.ld(col + 1);
      m_r0yz = rhoP.ld(col + sy + 1); m_c0yz = csP.ld(col + sy + 1);
      m_cp = PB(A.cp + rofs, pbytes).ld<CP>(col);
    }

    // ---- normal stresses (wgsl3d :1356-1453): x D+, y D-, z D-
    if (xp && ym && zm) {
      float dvx = 0.f, dvy = 0.f, dvz = 0.f;
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        dvx += dterm<FAST, 0>(A, c, wvx[ORD + c], wvx[ORD - 1 - c]);            // vx[x+c+1] - vx[x-c]
        if constexpr (PRE) {
          dvy += dterm<FAST, 1>(A, c, nvy_p[c], nvy_m[c]);
          dvz += dterm<FAST, 2>(A, c, nvz_p[c], nvz_m[c]);
        } else {
          dvy += dterm<FAST, 1>(A, c, vyP.ld(col + c * sy), vyP.ld(col - (c + 1) * sy));
          dvz += dterm<FAST, 2>(A, c, vzP.ld(col + c), vzP.ld(col - c - 1));
        }
      }
      if (bx >= 0) dvx = stretch<CP, FAST>(A.ax[0].ah, A.ax[0].bh, A.ax[0].kh, A.ax[0].khi, PB(A.psi[P_DVX_DX] + (long)bx * sx, pbytes), col, jx, dvx);
      if (by >= 0) dvy = stretch<CP, FAST>(A.ax[1].a, A.ax[1].b, A.ax[1].k, A.ax[1].ki, PB(A.psi[P_DVY_DY] + (long)x * nby * nz, pyb), py_col, jy, dvy);
      if (bz >= 0) dvz = stretch<CP, FAST>(A.ax[2].a, A.ax[2].b, A.ax[2].k, A.ax[2].ki, PB(A.psi[P_DVZ_DZ] + (long)x * ny * nbz, pzb), pz_col, jz, dvz);
      const float cp = PRE ? m_cp : PB(A.cp + rofs, pbytes).ld<CP>(col);
      const float lambda = rho0 * (cp * cp - 2.0f * cs0 * cs0);
      const float mu = rho0 * (cs0 * cs0);
      const float lp2m = lambda + 2.0f * mu;
      const PB sxxP(A.sxx + rofs, pbytes), syyP(A.syy + rofs, pbytes), szzP(A.szz + rofs, pbytes);
      if constexpr (VISCO && LC > 0) {
        float oxx[LC], oyy[LC], ozz[LC];
#pragma unroll
        for (int l = 0; l < LC; l++) {
          const long lo = (long)l * A.rs + rofs;
          oxx[l] = PB(A.r[0] + lo, pbytes).ld<CP>(col);
          oyy[l] = PB(A.r[1] + lo, pbytes).ld<CP>(col);
          ozz[l] = PB(A.r[2] + lo, pbytes).ld<CP>(col);
        }
        const float sxx0 = sxxP.ld<CP>(col), syy0 = syyP.ld<CP>(col), szz0 = szzP.ld<CP>(col);
        float Sxx = 0.f, Syy = 0.f, Szz = 0.f;
        const float div = dvx + dvy + dvz;
        const float div3 = FAST ? div * (1.0f / 3.0f) : div / 3.0f;
#pragma unroll
        for (int l = 0; l < LC; l++) {
          const float nxx = ((oxx[l] + div * A.zk.phi_p[l] - oxx[l] * A.zk.h_p[l]) * A.zk.m_p[l]);
          const float nyy = ((oyy[l] + (dvx - div3) * A.zk.phi_s[l] - oyy[l] * A.zk.h_s[l]) * A.zk.m_s[l]);
          const float nzz = ((ozz[l] + (dvy - div3) * A.zk.phi_s[l] - ozz[l] * A.zk.h_s[l]) * A.zk.m_s[l]);
          Sxx += oxx[l] + nxx; Syy += oyy[l] + nyy; Szz += ozz[l] + nzz;
          const long lo = (long)l * A.rs + rofs;
          PB(A.r[0] + lo, pbytes).st<CP>(col, nxx);
          PB(A.r[1] + lo, pbytes).st<CP>(col, nyy);
          PB(A.r[2] + lo, pbytes).st<CP>(col, nzz);
        }
        const float l23 = lambda + 2.0f * mu / 3.0f;
        sxxP.st<CP>(col, sxx0 + (lp2m * dvx + lambda * (dvy + dvz) + l23 * 0.5f * Sxx + mu * Syy) * dt);
        syyP.st<CP>(col, syy0 + (lambda * (dvx + dvz) + lp2m * dvy + l23 * 0.5f * Sxx + mu * Szz) * dt);
        szzP.st<CP>(col, szz0 + (lambda * (dvx + dvy) + lp2m * dvz + lp2m * 0.5f * Sxx - (mu / 3.0f) * (Syy + Szz)) * dt);
      } else if constexpr (VISCO) {
        float Sxx = 0.f, Syy = 0.f, Szz = 0.f;
        const float div = dvx + dvy + dvz;
        const float div3 = FAST ? div * (1.0f / 3.0f) : div / 3.0f;
        for (int l = 0; l < A.L; l++) {
          const long lo = (long)l * A.rs + rofs;
          const PB pxx(A.r[0] + lo, pbytes), pyy(A.r[1] + lo, pbytes), pzz(A.r[2] + lo, pbytes);
          const float oxx = pxx.ld<CP>(col), oyy = pyy.ld<CP>(col), ozz = pzz.ld<CP>(col);
          const float nxx = ((oxx + div * A.zk.phi_p[l] - oxx * A.zk.h_p[l]) * A.zk.m_p[l]);
          const float nyy = ((oyy + (dvx - div3) * A.zk.phi_s[l] - oyy * A.zk.h_s[l]) * A.zk.m_s[l]);
          const float nzz = ((ozz + (dvy - div3) * A.zk.phi_s[l] - ozz * A.zk.h_s[l]) * A.zk.m_s[l]);
          Sxx += oxx + nxx; Syy += oyy + nyy; Szz += ozz + nzz;
          pxx.st<CP>(col, nxx); pyy.st<CP>(col, nyy); pzz.st<CP>(col, nzz);
        }
        const float l23 = lambda + 2.0f * mu / 3.0f;
        sxxP.st<CP>(col, sxxP.ld<CP>(col) + (lp2m * dvx + lambda * (dvy + dvz) + l23 * 0.5f * Sxx + mu * Syy) * dt);
        syyP.st<CP>(col, syyP.ld<CP>(col) + (lambda * (dvx + dvz) + lp2m * dvy + l23 * 0.5f * Sxx + mu * Szz) * dt);
        szzP.st<CP>(col, szzP.ld<CP>(col) + (lambda * (dvx + dvy) + lp2m * dvz + lp2m * 0.5f * Sxx - (mu / 3.0f) * (Syy + Szz)) * dt);
      } else {
        sxxP.st<CP>(col, sxxP.ld<CP>(col) + (lp2m * dvx + lambda * (dvy + dvz)) * dt);
        syyP.st<CP>(col, syyP.ld<CP>(col) + (lambda * (dvx + dvz) + lp2m * dvy) * dt);
        szzP.st<CP>(col, szzP.ld<CP>(col) + (lambda * (dvx + dvy) + lp2m * dvz) * dt);
      }
    }

    // ---- shear stresses (wgsl3d :1457-1629)
#pragma unroll
    for (int sh = 0; sh < 3; sh++) {
      const bool ok = (sh < 2 ? xm : xp) && yp && zp;
      if (!ok) continue;
      float d1 = 0.f, d2 = 0.f, rho4, cs4;
      float *sig;
      float *rr;
      if (sh == 0) {            // sigma_xy: dvy/dx (D-), dvx/dy (D+)
#pragma unroll
        for (int c = 0; c < ORD; c++) {
          d1 += dterm<FAST, 0>(A, c, wvy[ORD + c], wvy[ORD - 1 - c]);         // vy[x+c] - vy[x-c-1]
          if constexpr (PRE) d2 += dterm<FAST, 1>(A, c, nxy_p[c], nxy_m[c]);
          else d2 += dterm<FAST, 1>(A, c, vxP.ld(col + (c + 1) * sy), vxP.ld(col - c * sy));
        }
        if (bx >= 0) d1 = stretch<CP, FAST>(A.ax[0].a, A.ax[0].b, A.ax[0].k, A.ax[0].ki, PB(A.psi[P_DVY_DX] + (long)bx * sx, pbytes), col, jx, d1);
        if (by >= 0) d2 = stretch<CP, FAST>(A.ax[1].ah, A.ax[1].bh, A.ax[1].kh, A.ax[1].khi, PB(A.psi[P_DVX_DY] + (long)x * nby * nz, pyb), py_col, jy, d2);
        if constexpr (PRE) {
          rho4 = 0.25f * (rho_n + rho0 + m_r0y + m_r1y);
          cs4 = 0.25f * (cs_n + cs0 + m_c0y + m_c1y);
        } else {
          rho4 = 0.25f * (rho_n + rho0 + rhoP.ld(col + sy) + rho1P.ld(col + sy));
          cs4 = 0.25f * (cs_n + cs0 + csP.ld(col + sy) + cs1P.ld(col + sy));
        }
        sig = A.sxy; rr = A.r[3];
      } else if (sh == 1) {     // sigma_xz: dvz/dx (D-), dvx/dz (D+)
#pragma unroll
        for (int c = 0; c < ORD; c++) {
          d1 += dterm<FAST, 0>(A, c, wvz[ORD + c], wvz[ORD - 1 - c]);         // vz[x+c] - vz[x-c-1]
          if constexpr (PRE) d2 += dterm<FAST, 2>(A, c, nxz_p[c], nxz_m[c]);
          else d2 += dterm<FAST, 2>(A, c, vxP.ld(col + c + 1), vxP.ld(col - c));
        }
        if (bx >= 0) d1 = stretch<CP, FAST>(A.ax[0].a, A.ax[0].b, A.ax[0].k, A.ax[0].ki, PB(A.psi[P_DVZ_DX] + (long)bx * sx, pbytes), col, jx, d1);
        if (bz >= 0) d2 = stretch<CP, FAST>(A.ax[2].ah, A.ax[2].bh, A.ax[2].kh, A.ax[2].khi, PB(A.psi[P_DVX_DZ] + (long)x * ny * nbz, pzb), pz_col, jz, d2);
        if constexpr (PRE) {
          rho4 = 0.25f * (rho_n + rho0 + m_r0z + m_r1z);
          cs4 = 0.25f * (cs_n + cs0 + m_c0z + m_c1z);
        } else {
          rho4 = 0.25f * (rho_n + rho0 + rhoP.ld(col + 1) + rho1P.ld(col + 1));
          cs4 = 0.25f * (cs_n + cs0 + csP.ld(col + 1) + cs1P.ld(col + 1));
        }
        sig = A.sxz; rr = A.r[4];
      } else {                  // sigma_yz: dvz/dy (D+), dvy/dz (D+)
#pragma unroll
        for (int c = 0; c < ORD; c++) {
          if constexpr (PRE) {
            d1 += dterm<FAST, 1>(A, c, nzy_p[c], nzy_m[c]);
            d2 += dterm<FAST, 2>(A, c, nyz_p[c], nyz_m[c]);
          } else {
            d1 += dterm<FAST, 1>(A, c, vzP.ld(col + (c + 1) * sy), vzP.ld(col - c * sy));
            d2 += dterm<FAST, 2>(A, c, vyP.ld(col + c + 1), vyP.ld(col - c));
          }
        }
        if (by >= 0) d1 = stretch<CP, FAST>(A.ax[1].ah, A.ax[1].bh, A.ax[1].kh, A.ax[1].khi, PB(A.psi[P_DVZ_DY] + (long)x * nby * nz, pyb), py_col, jy, d1);
        if (bz >= 0) d2 = stretch<CP, FAST>(A.ax[2].ah, A.ax[2].bh, A.ax[2].kh, A.ax[2].khi, PB(A.psi[P_DVY_DZ] + (long)x * ny * nbz, pzb), pz_col, jz, d2);
        if constexpr (PRE) {
          rho4 = 0.25f * (m_r0y + rho0 + m_r0z + m_r0yz);
          cs4 = 0.25f * (m_c0y + cs0 + m_c0z + m_c0yz);
        } else {
          rho4 = 0.25f * (rhoP.ld(col + sy) + rho0 + rhoP.ld(col + 1) + rhoP.ld(col + sy + 1));
          cs4 = 0.25f * (csP.ld(col + sy) + cs0 + csP.ld(col + 1) + csP.ld(col + sy + 1));
        }
        sig = A.syz; rr = A.r[5];
      }
      const float mu = rho4 * (cs4 * cs4);
      const PB sigP(sig + rofs, pbytes);
      if constexpr (VISCO && LC > 0) {
        float o[LC];
#pragma unroll
        for (int l = 0; l < LC; l++) o[l] = PB(rr + (long)l * A.rs + rofs, pbytes).ld<CP>(col);
        const float sig0 = sigP.ld<CP>(col);
        float S = 0.f;
        const float X = d1 + d2;
#pragma unroll
        for (int l = 0; l < LC; l++) {
          const float n = (o[l] + X * A.zk.phi_s[l] - o[l] * A.zk.h_s[l]) * A.zk.m_s[l];
          S += o[l] + n;
          PB(rr + (long)l * A.rs + rofs, pbytes).st<CP>(col, n);
        }
        sigP.st<CP>(col, sig0 + (mu * (d1 + d2) + mu * 0.5f * S) * dt);
      } else if constexpr (VISCO) {
        float S = 0.f;
        const float X = d1 + d2;
        for (int l = 0; l < A.L; l++) {
          const PB pr(rr + (long)l * A.rs + rofs, pbytes);
          const float o = pr.ld<CP>(col);
          const float n = (o + X * A.zk.phi_s[l] - o * A.zk.h_s[l]) * A.zk.m_s[l];
          S += o + n;
          pr.st<CP>(col, n);
        }
        sigP.st<CP>(col, sigP.ld<CP>(col) + (mu * (d1 + d2) + mu * 0.5f * S) * dt);
      } else {
        sigP.st<CP>(col, sigP.ld<CP>(col) + (d2 + d1) * mu * dt);
      }
    }
  }
}

template <int ORD, int TY, bool NT, bool FAST>
__global__ __launch_bounds__(BZ * TY) void k_velocity3d(StepArgs A, int xchunk, int nzb, int nyb, int nxc, int remap) {
  constexpr int CP = NT ? 2 : 0;
  const Tile T = tile_of_block(nzb, nyb, nxc, remap);
  const int z = T.bz * BZ + threadIdx.x;
  const int y = T.by * TY + threadIdx.y;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  const bool valid = z < nz && y < ny;
  constexpr int last = ORD - 1, off = ORD - 1;
  const unsigned sy = nz;
  const long sx = (long)ny * nz;
  const unsigned pbytes = (unsigned)(sx * 4);
  const unsigned col = y * sy + z;
  const float dt = A.dt;
  const int x0 = T.bxc * xchunk, x1 = min(nx, x0 + xchunk);
  // whole-field buffers (the host launches this kernel for fewer than 2^30 cells)
  const FBuf Fsxx(A.sxx, A.ncell), Fsxy(A.sxy, A.ncell), Fsxz(A.sxz, A.ncell), Fsyy(A.syy, A.ncell);
  const FBuf Fsyz(A.syz, A.ncell), Fszz(A.szz, A.ncell), Frho(A.rho, A.ncell);
  const FBuf Fvx(A.vx, A.ncell), Fvy(A.vy, A.ncell), Fvz(A.vz, A.ncell);
  // first element of plane x, clamped into the grid (clamped planes feed no update)
  auto pl = [&](int x) { return (unsigned)((x < 0 ? 0 : (x >= nx ? nx - 1 : x)) * sx); };
  unsigned v2bits = 0;
  if (valid) {
    const int jy = y - off, jz = z - off;
    const int by = band_index(A.ax[1], jy), bz = band_index(A.ax[2], jz);
    const int nby = A.ax[1].nb, nbz = A.ax[2].nb;
    const unsigned py_col = by * nz + z, pz_col = y * nbz + bz;
    const unsigned pyb = (unsigned)nby * nz * 4u, pzb = (unsigned)ny * nbz * 4u;
    const bool yi_vx = y >= last + 1 && y < ny - last, zi_vx = z >= last + 1 && z < nz - last;
    const bool yi_vy = y >= last && y < ny - last - 1, zi_vy = z >= last + 1 && z < nz - last;
    const bool yi_vz = y >= last + 1 && y < ny - last, zi_vz = z >= last && z < nz - last - 1;
    // x windows: sxx at x-ORD..x+ORD-1 (D-), sxy/sxz at x-(ORD-1)..x+ORD (D+)
    float wxx[2 * ORD], wxy[2 * ORD], wxz[2 * ORD];
#pragma unroll
    for (int i = 0; i < 2 * ORD - 1; i++) {
      wxx[i + 1] = Fsxx.ld(pl(x0 - ORD + i) + col);
      wxy[i + 1] = Fsxy.ld(pl(x0 - (ORD - 1) + i) + col);
      wxz[i + 1] = Fsxz.ld(pl(x0 - (ORD - 1) + i) + col);
    }
    float rho_n = Frho.ld(pl(x0) + col);
#pragma unroll 1
    for (int x = x0; x < x1; x++) {
#pragma unroll
      for (int i = 0; i < 2 * ORD - 1; i++) { wxx[i] = wxx[i + 1]; wxy[i] = wxy[i + 1]; wxz[i] = wxz[i + 1]; }
      wxx[2 * ORD - 1] = Fsxx.ld(pl(x + ORD - 1) + col);
      wxy[2 * ORD - 1] = Fsxy.ld(pl(x + ORD) + col);
      wxz[2 * ORD - 1] = Fsxz.ld(pl(x + ORD) + col);
      const float rho0 = rho_n;
      rho_n = Frho.ld(pl(x + 1) + col);
      const int jx = x - off;
      const int bx = band_index(A.ax[0], jx);
      const long rofs = (long)x * sx;
      const unsigned i0 = (unsigned)rofs + col;      // the cell in plane x
      // Every load of the plane first, unconditionally: a neighbour index
      // outside the plane reads the adjacent plane, or 0 outside the field,
      // and such values are used only outside the interiors, where they are
      // not; so the wavefront has all of them in flight before any arithmetic.
      float vxn = Fvx.ld<CP>(i0), vyn = Fvy.ld<CP>(i0), vzn = Fvz.ld<CP>(i0);
      float xy_p[ORD], xy_m[ORD], xz_p[ORD], xz_m[ORD];      // vx: sxy along y, sxz along z (D-)
      float yy_p[ORD], yy_m[ORD], yz_p[ORD], yz_m[ORD];      // vy: syy along y (D+), syz along z (D-)
      float zy_p[ORD], zy_m[ORD], zz_p[ORD], zz_m[ORD];      // vz: syz along y (D-), szz along z (D+)
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        // the own cell of plane x is already in the x windows
        xy_p[c] = c == 0 ? wxy[ORD - 1] : Fsxy.ld(i0 + c * sy); xy_m[c] = Fsxy.ld(i0 - (c + 1) * sy);
        xz_p[c] = c == 0 ? wxz[ORD - 1] : Fsxz.ld(i0 + c); xz_m[c] = Fsxz.ld(i0 - c - 1);
        yy_p[c] = Fsyy.ld(i0 + (c + 1) * sy); yy_m[c] = Fsyy.ld(i0 - c * sy);
        yz_p[c] = Fsyz.ld(i0 + c); yz_m[c] = Fsyz.ld(i0 - c - 1);
        zy_p[c] = Fsyz.ld(i0 + c * sy); zy_m[c] = Fsyz.ld(i0 - (c + 1) * sy);
        zz_p[c] = Fszz.ld(i0 + c + 1); zz_m[c] = Fszz.ld(i0 - c);
      }
      const float rho_y1 = Frho.ld(i0 + sy);
      // vx (wgsl3d :1647-1681)
      if (x >= last + 1 && x < nx - last && yi_vx && zi_vx) {
        float a = 0.f, b = 0.f, c_ = 0.f;
#pragma unroll
        for (int c = 0; c < ORD; c++) {
          a += dterm<FAST, 0>(A, c, wxx[ORD + c], wxx[ORD - 1 - c]);          // sxx[x+c] - sxx[x-c-1]
          b += dterm<FAST, 1>(A, c, xy_p[c], xy_m[c]);
          c_ += dterm<FAST, 2>(A, c, xz_p[c], xz_m[c]);
        }
        if (bx >= 0) a = stretch<CP, FAST>(A.ax[0].a, A.ax[0].b, A.ax[0].k, A.ax[0].ki, PB(A.psi[P_DSXX_DX] + (long)bx * sx, pbytes), col, jx, a);
        if (by >= 0) b = stretch<CP, FAST>(A.ax[1].a, A.ax[1].b, A.ax[1].k, A.ax[1].ki, PB(A.psi[P_DSXY_DY] + (long)x * nby * nz, pyb), py_col, jy, b);
        if (bz >= 0) c_ = stretch<CP, FAST>(A.ax[2].a, A.ax[2].b, A.ax[2].k, A.ax[2].ki, PB(A.psi[P_DSXZ_DZ] + (long)x * ny * nbz, pzb), pz_col, jz, c_);
        const float rho = 0.5f * (rho_n + rho0);
        if (rho > 0.f) vxn = (a + b + c_) * dt / rho + vxn;
      }
      // vy (wgsl3d :1683-1717)
      const bool xi_vyz = x >= last && x < nx - last - 1;
      if (xi_vyz && yi_vy && zi_vy) {
        float a = 0.f, b = 0.f, c_ = 0.f;
#pragma unroll
        for (int c = 0; c < ORD; c++) {
          a += dterm<FAST, 0>(A, c, wxy[ORD + c], wxy[ORD - 1 - c]);          // sxy[x+c+1] - sxy[x-c]
          b += dterm<FAST, 1>(A, c, yy_p[c], yy_m[c]);
          c_ += dterm<FAST, 2>(A, c, yz_p[c], yz_m[c]);
        }
        if (bx >= 0) a = stretch<CP, FAST>(A.ax[0].ah, A.ax[0].bh, A.ax[0].kh, A.ax[0].khi, PB(A.psi[P_DSXY_DX] + (long)bx * sx, pbytes), col, jx, a);
        if (by >= 0) b = stretch<CP, FAST>(A.ax[1].ah, A.ax[1].bh, A.ax[1].kh, A.ax[1].khi, PB(A.psi[P_DSYY_DY] + (long)x * nby * nz, pyb), py_col, jy, b);
        if (bz >= 0) c_ = stretch<CP, FAST>(A.ax[2].a, A.ax[2].b, A.ax[2].k, A.ax[2].ki, PB(A.psi[P_DSYZ_DZ] + (long)x * ny * nbz, pzb), pz_col, jz, c_);
        const float rho = 0.5f * (rho_y1 + rho0);
        if (rho > 0.f) vyn = (a + b + c_) * dt / rho + vyn;
      }
      // vz, with the vy density (wgsl3d :1719-1758)
      if (xi_vyz && yi_vz && zi_vz) {
        float a = 0.f, b = 0.f, c_ = 0.f;
#pragma unroll
        for (int c = 0; c < ORD; c++) {
          a += dterm<FAST, 0>(A, c, wxz[ORD + c], wxz[ORD - 1 - c]);          // sxz[x+c+1] - sxz[x-c]
          b += dterm<FAST, 1>(A, c, zy_p[c], zy_m[c]);
          c_ += dterm<FAST, 2>(A, c, zz_p[c], zz_m[c]);
        }
        if (bx >= 0) a = stretch<CP, FAST>(A.ax[0].ah, A.ax[0].bh, A.ax[0].kh, A.ax[0].khi, PB(A.psi[P_DSXZ_DX] + (long)bx * sx, pbytes), col, jx, a);
        if (by >= 0) b = stretch<CP, FAST>(A.ax[1].a, A.ax[1].b, A.ax[1].k, A.ax[1].ki, PB(A.psi[P_DSYZ_DY] + (long)x * nby * nz, pyb), py_col, jy, b);
        if (bz >= 0) c_ = stretch<CP, FAST>(A.ax[2].ah, A.ax[2].bh, A.ax[2].kh, A.ax[2].khi, PB(A.psi[P_DSZZ_DZ] + (long)x * ny * nbz, pzb), pz_col, jz, c_);
        const float rho = 0.5f * (rho_y1 + rho0);
        if (rho > 0.f) vzn = (a + b + c_) * dt / rho + vzn;
      } else {
        vxn = 0.f; vyn = 0.f; vzn = 0.f;
      }
      Fvx.st<CP>(i0, vxn);
      Fvy.st<CP>(i0, vyn);
      Fvz.st<CP>(i0, vzn);
      const unsigned u = __float_as_uint(vxn * vxn + vyn * vyn + vzn * vzn);
      v2bits = v2bits > u ? v2bits : u;
    }
  }
  block_max_u32(v2bits, A.v2max);
}

// The same velocity pass with the y-neighbours of sigma_xy, sigma_yy and
// sigma_yz staged in LDS: per x-plane, each thread puts its own cell of the
// three fields into the tile's rows, the first 2*ORD wavefronts load the ORD
// halo rows above and below the tile, and after one barrier every y-stencil
// operand comes from LDS.  A tile then loads TY + 2*ORD rows of those fields
// per plane instead of TY * (2*ORD + ...) separate row loads that mostly hit
// other tiles' rows.  The LDS rows are double-buffered by plane parity, so a
// plane needs one barrier: a wavefront writes buffer p only after every
// wavefront has passed the barrier of the plane that read p before.  Same
// operands, same arithmetic: bit-identical to k_velocity3d.  Needs TY >= 2*ORD.
template <int ORD, int TY, bool NT, bool FAST>
__global__ __launch_bounds__(BZ * TY) void k_velocity3d_lds(StepArgs A, int xchunk, int nzb, int nyb, int nxc, int remap) {
  static_assert(TY >= 2 * ORD, "the halo rows are loaded by the first 2*ORD wavefronts");
  constexpr int CP = NT ? 2 : 0;
  constexpr int ROWS = TY + 2 * ORD;
  __shared__ float s_xy[2][ROWS][BZ], s_yy[2][ROWS][BZ], s_yz[2][ROWS][BZ];
  const Tile T = tile_of_block(nzb, nyb, nxc, remap);
  const int ty = threadIdx.y;
  const int z = T.bz * BZ + threadIdx.x;
  const int y0 = T.by * TY;
  const int y = y0 + ty;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  const bool valid = z < nz && y < ny;
  constexpr int last = ORD - 1, off = ORD - 1;
  const unsigned sy = nz;
  const long sx = (long)ny * nz;
  const unsigned pbytes = (unsigned)(sx * 4);
  const float dt = A.dt;
  const int x0 = T.bxc * xchunk, x1 = min(nx, x0 + xchunk);
  const FBuf Fsxx(A.sxx, A.ncell), Fsxy(A.sxy, A.ncell), Fsxz(A.sxz, A.ncell), Fsyy(A.syy, A.ncell);
  const FBuf Fsyz(A.syz, A.ncell), Fszz(A.szz, A.ncell), Frho(A.rho, A.ncell);
  const FBuf Fvx(A.vx, A.ncell), Fvy(A.vy, A.ncell), Fvz(A.vz, A.ncell);
  auto pl = [&](int x) { return (unsigned)((x < 0 ? 0 : (x >= nx ? nx - 1 : x)) * sx); };
  // column of this thread (z clamped for the lanes past nz, which compute nothing)
  const unsigned col = (unsigned)(min(y, ny - 1)) * sy + (unsigned)min(z, nz - 1);
  // halo row loaded by this wavefront: ORD rows above the tile, then ORD below
  const bool halo = ty < 2 * ORD;
  const int hrow = ty < ORD ? ty : TY + ty;                    // LDS row
  const int hy = y0 - ORD + hrow;                              // grid row (may lie outside [0, ny))
  const unsigned hcol = (unsigned)((long)hy * sy + min(z, nz - 1));
  const int jy = y - off, jz = z - off;
  const int by = band_index(A.ax[1], jy), bz = band_index(A.ax[2], jz);
  const int nby = A.ax[1].nb, nbz = A.ax[2].nb;
  const unsigned py_col = by * nz + z, pz_col = y * nbz + bz;
  const unsigned pyb = (unsigned)nby * nz * 4u, pzb = (unsigned)ny * nbz * 4u;
  const bool yi_vx = y >= last + 1 && y < ny - last, zi_vx = z >= last + 1 && z < nz - last;
  const bool yi_vy = y >= last && y < ny - last - 1, zi_vy = z >= last + 1 && z < nz - last;
  const bool yi_vz = y >= last + 1 && y < ny - last, zi_vz = z >= last && z < nz - last - 1;
  float wxx[2 * ORD], wxy[2 * ORD], wxz[2 * ORD];
#pragma unroll
  for (int i = 0; i < 2 * ORD - 1; i++) {
    wxx[i + 1] = Fsxx.ld(pl(x0 - ORD + i) + col);
    wxy[i + 1] = Fsxy.ld(pl(x0 - (ORD - 1) + i) + col);
    wxz[i + 1] = Fsxz.ld(pl(x0 - (ORD - 1) + i) + col);
  }
  float rho_n = Frho.ld(pl(x0) + col);
  unsigned v2bits = 0;
#pragma unroll 1
  for (int x = x0; x < x1; x++) {
    const int p = x & 1;
#pragma unroll
    for (int i = 0; i < 2 * ORD - 1; i++) { wxx[i] = wxx[i + 1]; wxy[i] = wxy[i + 1]; wxz[i] = wxz[i + 1]; }
    wxx[2 * ORD - 1] = Fsxx.ld(pl(x + ORD - 1) + col);
    wxy[2 * ORD - 1] = Fsxy.ld(pl(x + ORD) + col);
    wxz[2 * ORD - 1] = Fsxz.ld(pl(x + ORD) + col);
    const float rho0 = rho_n;
    rho_n = Frho.ld(pl(x + 1) + col);
    const long rofs = (long)x * sx;
    const unsigned i0 = (unsigned)rofs + col;
    // own cells of the three y-stencil fields into the tile rows, halo rows
    // from the neighbouring tiles (a row outside the grid reads the adjacent
    // plane or 0, used only outside the interiors, as in k_velocity3d)
    s_xy[p][ty + ORD][threadIdx.x] = wxy[ORD - 1];
    s_yy[p][ty + ORD][threadIdx.x] = Fsyy.ld(i0);
    s_yz[p][ty + ORD][threadIdx.x] = Fsyz.ld(i0);
    if (halo) {
      const unsigned h0 = (unsigned)rofs + hcol;
      s_xy[p][hrow][threadIdx.x] = Fsxy.ld(h0);
      s_yy[p][hrow][threadIdx.x] = Fsyy.ld(h0);
      s_yz[p][hrow][threadIdx.x] = Fsyz.ld(h0);
    }
    float vxn = Fvx.ld<CP>(i0), vyn = Fvy.ld<CP>(i0), vzn = Fvz.ld<CP>(i0);
    float xz_p[ORD], xz_m[ORD], yz_p[ORD], yz_m[ORD], zz_p[ORD], zz_m[ORD];
#pragma unroll
    for (int c = 0; c < ORD; c++) {
      xz_p[c] = c == 0 ? wxz[ORD - 1] : Fsxz.ld(i0 + c); xz_m[c] = Fsxz.ld(i0 - c - 1);
      yz_p[c] = Fsyz.ld(i0 + c); yz_m[c] = Fsyz.ld(i0 - c - 1);
      zz_p[c] = Fszz.ld(i0 + c + 1); zz_m[c] = Fszz.ld(i0 - c);
    }
    const float rho_y1 = Frho.ld(i0 + sy);
    __syncthreads();
    float xy_p[ORD], xy_m[ORD], yy_p[ORD], yy_m[ORD], zy_p[ORD], zy_m[ORD];
#pragma unroll
    for (int c = 0; c < ORD; c++) {
      xy_p[c] = s_xy[p][ty + ORD + c][threadIdx.x]; xy_m[c] = s_xy[p][ty + ORD - c - 1][threadIdx.x];
      yy_p[c] = s_yy[p][ty + ORD + c + 1][threadIdx.x]; yy_m[c] = s_yy[p][ty + ORD - c][threadIdx.x];
      zy_p[c] = s_yz[p][ty + ORD + c][threadIdx.x]; zy_m[c] = s_yz[p][ty + ORD - c - 1][threadIdx.x];
    }
    if (!valid) continue;
    const int jx = x - off;
    const int bx = band_index(A.ax[0], jx);
    // vx (wgsl3d :1647-1681)
    if (x >= last + 1 && x < nx - last && yi_vx && zi_vx) {
      float a = 0.f, b = 0.f, c_ = 0.f;
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        a += dterm<FAST, 0>(A, c, wxx[ORD + c], wxx[ORD - 1 - c]);
        b += dterm<FAST, 1>(A, c, xy_p[c], xy_m[c]);
        c_ += dterm<FAST, 2>(A, c, xz_p[c], xz_m[c]);
      }
      if (bx >= 0) a = stretch<CP, FAST>(A.ax[0].a, A.ax[0].b, A.ax[0].k, A.ax[0].ki, PB(A.psi[P_DSXX_DX] + (long)bx * sx, pbytes), col, jx, a);
      if (by >= 0) b = stretch<CP, FAST>(A.ax[1].a, A.ax[1].b, A.ax[1].k, A.ax[1].ki, PB(A.psi[P_DSXY_DY] + (long)x * nby * nz, pyb), py_col, jy, b);
      if (bz >= 0) c_ = stretch<CP, FAST>(A.ax[2].a, A.ax[2].b, A.ax[2].k, A.ax[2].ki, PB(A.psi[P_DSXZ_DZ] + (long)x * ny * nbz, pzb), pz_col, jz, c_);
      const float rho = 0.5f * (rho_n + rho0);
      if (rho > 0.f) vxn = (a + b + c_) * dt / rho + vxn;
    }
    // vy (wgsl3d :1683-1717)
    const bool xi_vyz = x >= last && x < nx - last - 1;
    if (xi_vyz && yi_vy && zi_vy) {
      float a = 0.f, b = 0.f, c_ = 0.f;
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        a += dterm<FAST, 0>(A, c, wxy[ORD + c], wxy[ORD - 1 - c]);
        b += dterm<FAST, 1>(A, c, yy_p[c], yy_m[c]);
        c_ += dterm<FAST, 2>(A, c, yz_p[c], yz_m[c]);
      }
      if (bx >= 0) a = stretch<CP, FAST>(A.ax[0].ah, A.ax[0].bh, A.ax[0].kh, A.ax[0].khi, PB(A.psi[P_DSXY_DX] + (long)bx * sx, pbytes), col, jx, a);
      if (by >= 0) b = stretch<CP, FAST>(A.ax[1].ah, A.ax[1].bh, A.ax[1].kh, A.ax[1].khi, PB(A.psi[P_DSYY_DY] + (long)x * nby * nz, pyb), py_col, jy, b);
      if (bz >= 0) c_ = stretch<CP, FAST>(A.ax[2].a, A.ax[2].b, A.ax[2].k, A.ax[2].ki, PB(A.psi[P_DSYZ_DZ] + (long)x * ny * nbz, pzb), pz_col, jz, c_);
      const float rho = 0.5f * (rho_y1 + rho0);
      if (rho > 0.f) vyn = (a + b + c_) * dt / rho + vyn;
    }
    // vz, with the vy density (wgsl3d :1719-1758)
    if (xi_vyz && yi_vz && zi_vz) {
      float a = 0.f, b = 0.f, c_ = 0.f;
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        a += dterm<FAST, 0>(A, c, wxz[ORD + c], wxz[ORD - 1 - c]);
        b += dterm<FAST, 1>(A, c, zy_p[c], zy_m[c]);
        c_ += dterm<FAST, 2>(A, c, zz_p[c], zz_m[c]);
      }
      if (bx >= 0) a = stretch<CP, FAST>(A.ax[0].ah, A.ax[0].bh, A.ax[0].kh, A.ax[0].khi, PB(A.psi[P_DSXZ_DX] + (long)bx * sx, pbytes), col, jx, a);
      if (by >= 0) b = stretch<CP, FAST>(A.ax[1].a, A.ax[1].b, A.ax[1].k, A.ax[1].ki, PB(A.psi[P_DSYZ_DY] + (long)x * nby * nz, pyb), py_col, jy, b);
      if (bz >= 0) c_ = stretch<CP, FAST>(A.ax[2].ah, A.ax[2].bh, A.ax[2].kh, A.ax[2].khi, PB(A.psi[P_DSZZ_DZ] + (long)x * ny * nbz, pzb), pz_col, jz, c_);
      const float rho = 0.5f * (rho_y1 + rho0);
      if (rho > 0.f) vzn = (a + b + c_) * dt / rho + vzn;
    } else {
      vxn = 0.f; vyn = 0.f; vzn = 0.f;
    }
    Fvx.st<CP>(i0, vxn);
    Fvy.st<CP>(i0, vyn);
    Fvz.st<CP>(i0, vzn);
    const unsigned u = __float_as_uint(vxn * vxn + vyn * vyn + vzn * vzn);
    v2bits = v2bits > u ? v2bits : u;
  }
  block_max_u32(v2bits, A.v2max);
}

// ---------------------------------------------------------------------------
// Plain kernels: one cell per thread, grid (z-tiles, y-tiles, x), flat global
// loads; neighbours in x come from L2/MALL.  The autotuner's xchunk = 0.
constexpr int BYY = 4;

template <bool FAST>
__device__ __forceinline__ float cpml_stretch_m(const float *a, const float *b, const float *k, const float *ki,
                                                float *psi, long pidx, int j, bool in_band, float D) {
  if (!in_band) return D;
  const float pn = b[j] * psi[pidx] + a[j] * D;
  psi[pidx] = pn;
  if constexpr (FAST) return D * ki[j] + pn;
  else return D / k[j] + pn;
}

template <bool VISCO, int ORD, bool FAST, int LC = 0>
__global__ __launch_bounds__(BZ * BYY) void k_stress3d_plain(StepArgs A) {
  const int z = blockIdx.x * BZ + threadIdx.x;
  const int y = blockIdx.y * BYY + threadIdx.y;
  const int x = blockIdx.z;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  if (z >= nz || y >= ny) return;
  constexpr int last = ORD - 1, off = ORD - 1;
  const long sy = nz, sx = (long)ny * nz;
  const long id = x * sx + y * sy + z;
  const float dt = A.dt;
  const int jx = x - off, jy = y - off, jz = z - off;
  const int bx = band_index(A.ax[0], jx), by = band_index(A.ax[1], jy), bz = band_index(A.ax[2], jz);
  const int nbx = A.ax[0].nb, nby = A.ax[1].nb, nbz = A.ax[2].nb;
  // compact psi indices for x-, y- and z-derivatives at this cell
  const long px = (long)bx * sx + y * sy + z;
  const long py = (long)x * nby * nz + (long)by * nz + z;
  const long pz = ((long)x * ny + y) * nbz + bz;
  (void)nbx;

  // ---- normal stresses (wgsl3d :1356-1453)
  if (x >= last && x < nx - last - 1 && y >= last + 1 && y < ny - last && z >= last + 1 && z < nz - last) {
    float dvx = 0.f, dvy = 0.f, dvz = 0.f;
#pragma unroll
    for (int c = 0; c < ORD; c++) {
      dvx += dterm<FAST, 0>(A, c, A.vx[id + (c + 1) * sx], A.vx[id - c * sx]);
      dvy += dterm<FAST, 1>(A, c, A.vy[id + c * sy], A.vy[id - (c + 1) * sy]);
      dvz += dterm<FAST, 2>(A, c, A.vz[id + c], A.vz[id - c - 1]);
    }
    dvx = cpml_stretch_m<FAST>(A.ax[0].ah, A.ax[0].bh, A.ax[0].kh, A.ax[0].khi, A.psi[P_DVX_DX], px, jx, bx >= 0, dvx);
    dvy = cpml_stretch_m<FAST>(A.ax[1].a, A.ax[1].b, A.ax[1].k, A.ax[1].ki, A.psi[P_DVY_DY], py, jy, by >= 0, dvy);
    dvz = cpml_stretch_m<FAST>(A.ax[2].a, A.ax[2].b, A.ax[2].k, A.ax[2].ki, A.psi[P_DVZ_DZ], pz, jz, bz >= 0, dvz);
    const float rho = A.rho[id], cp = A.cp[id], cs = A.cs[id];
    const float lambda = rho * (cp * cp - 2.0f * cs * cs);
    const float mu = rho * (cs * cs);
    const float lp2m = lambda + 2.0f * mu;
    if constexpr (VISCO && LC > 0) {      // all own-cell loads first (see k_stress3d)
      float oxx[LC], oyy[LC], ozz[LC];
#pragma unroll
      for (int l = 0; l < LC; l++) {
        oxx[l] = A.r[0][l * A.rs + id];
        oyy[l] = A.r[1][l * A.rs + id];
        ozz[l] = A.r[2][l * A.rs + id];
      }
      const float sxx0 = A.sxx[id], syy0 = A.syy[id], szz0 = A.szz[id];
      float Sxx = 0.f, Syy = 0.f, Szz = 0.f;
      const float div = dvx + dvy + dvz;
      const float div3 = FAST ? div * (1.0f / 3.0f) : div / 3.0f;
#pragma unroll
      for (int l = 0; l < LC; l++) {
        const float nxx = ((oxx[l] + div * A.zk.phi_p[l] - oxx[l] * A.zk.h_p[l]) * A.zk.m_p[l]);
        const float nyy = ((oyy[l] + (dvx - div3) * A.zk.phi_s[l] - oyy[l] * A.zk.h_s[l]) * A.zk.m_s[l]);
        const float nzz = ((ozz[l] + (dvy - div3) * A.zk.phi_s[l] - ozz[l] * A.zk.h_s[l]) * A.zk.m_s[l]);
        Sxx += oxx[l] + nxx; Syy += oyy[l] + nyy; Szz += ozz[l] + nzz;
        A.r[0][l * A.rs + id] = nxx;
        A.r[1][l * A.rs + id] = nyy;
        A.r[2][l * A.rs + id] = nzz;
      }
      const float l23 = lambda + 2.0f * mu / 3.0f;
      A.sxx[id] = sxx0 + (lp2m * dvx + lambda * (dvy + dvz) + l23 * 0.5f * Sxx + mu * Syy) * dt;
      A.syy[id] = syy0 + (lambda * (dvx + dvz) + lp2m * dvy + l23 * 0.5f * Sxx + mu * Szz) * dt;
      A.szz[id] = szz0 + (lambda * (dvx + dvy) + lp2m * dvz + lp2m * 0.5f * Sxx - (mu / 3.0f) * (Syy + Szz)) * dt;
    } else if constexpr (VISCO) {
      float Sxx = 0.f, Syy = 0.f, Szz = 0.f;
      const float div = dvx + dvy + dvz;
      const float div3 = FAST ? div * (1.0f / 3.0f) : div / 3.0f;
      for (int l = 0; l < A.L; l++) {
        float *pxx = A.r[0] + l * A.rs + id, *pyy = A.r[1] + l * A.rs + id, *pzz = A.r[2] + l * A.rs + id;
        const float oxx = *pxx, oyy = *pyy, ozz = *pzz;
        const float nxx = ((oxx + div * A.zk.phi_p[l] - oxx * A.zk.h_p[l]) * A.zk.m_p[l]);
        const float nyy = ((oyy + (dvx - div3) * A.zk.phi_s[l] - oyy * A.zk.h_s[l]) * A.zk.m_s[l]);
        const float nzz = ((ozz + (dvy - div3) * A.zk.phi_s[l] - ozz * A.zk.h_s[l]) * A.zk.m_s[l]);
        Sxx += oxx + nxx; Syy += oyy + nyy; Szz += ozz + nzz;
        *pxx = nxx; *pyy = nyy; *pzz = nzz;
      }
      const float l23 = lambda + 2.0f * mu / 3.0f;
      A.sxx[id] = A.sxx[id] + (lp2m * dvx + lambda * (dvy + dvz) + l23 * 0.5f * Sxx + mu * Syy) * dt;
      A.syy[id] = A.syy[id] + (lambda * (dvx + dvz) + lp2m * dvy + l23 * 0.5f * Sxx + mu * Szz) * dt;
      A.szz[id] = A.szz[id] + (lambda * (dvx + dvy) + lp2m * dvz + lp2m * 0.5f * Sxx - (mu / 3.0f) * (Syy + Szz)) * dt;
    } else {
      A.sxx[id] = A.sxx[id] + (lp2m * dvx + lambda * (dvy + dvz)) * dt;
      A.syy[id] = A.syy[id] + (lambda * (dvx + dvz) + lp2m * dvy) * dt;
      A.szz[id] = A.szz[id] + (lambda * (dvx + dvy) + lp2m * dvz) * dt;
    }
  }

  const bool xm = x >= last + 1 && x < nx - last;      // D- range in x
  const bool xp = x >= last && x < nx - last - 1;      // D+ range in x
  const bool yp = y >= last && y < ny - last - 1;
  const bool zp = z >= last && z < nz - last - 1;
  // ---- shear stresses: sigma_xy, sigma_xz, sigma_yz (wgsl3d :1457-1629)
#pragma unroll
  for (int sh = 0; sh < 3; sh++) {
    const bool ok = (sh < 2 ? xm : xp) && yp && zp;
    if (!ok) continue;
    float d1 = 0.f, d2 = 0.f, rho4, cs4;
    float *sig, *rr;
    if (sh == 0) {            // dvy/dx (D-), dvx/dy (D+)
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        d1 += dterm<FAST, 0>(A, c, A.vy[id + c * sx], A.vy[id - (c + 1) * sx]);
        d2 += dterm<FAST, 1>(A, c, A.vx[id + (c + 1) * sy], A.vx[id - c * sy]);
      }
      d1 = cpml_stretch_m<FAST>(A.ax[0].a, A.ax[0].b, A.ax[0].k, A.ax[0].ki, A.psi[P_DVY_DX], px, jx, bx >= 0, d1);
      d2 = cpml_stretch_m<FAST>(A.ax[1].ah, A.ax[1].bh, A.ax[1].kh, A.ax[1].khi, A.psi[P_DVX_DY], py, jy, by >= 0, d2);
      rho4 = 0.25f * (A.rho[id + sx] + A.rho[id] + A.rho[id + sy] + A.rho[id + sx + sy]);
      cs4 = 0.25f * (A.cs[id + sx] + A.cs[id] + A.cs[id + sy] + A.cs[id + sx + sy]);
      sig = A.sxy; rr = A.r[3];
    } else if (sh == 1) {     // dvz/dx (D-), dvx/dz (D+)
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        d1 += dterm<FAST, 0>(A, c, A.vz[id + c * sx], A.vz[id - (c + 1) * sx]);
        d2 += dterm<FAST, 2>(A, c, A.vx[id + c + 1], A.vx[id - c]);
      }
      d1 = cpml_stretch_m<FAST>(A.ax[0].a, A.ax[0].b, A.ax[0].k, A.ax[0].ki, A.psi[P_DVZ_DX], px, jx, bx >= 0, d1);
      d2 = cpml_stretch_m<FAST>(A.ax[2].ah, A.ax[2].bh, A.ax[2].kh, A.ax[2].khi, A.psi[P_DVX_DZ], pz, jz, bz >= 0, d2);
      rho4 = 0.25f * (A.rho[id + sx] + A.rho[id] + A.rho[id + 1] + A.rho[id + sx + 1]);
      cs4 = 0.25f * (A.cs[id + sx] + A.cs[id] + A.cs[id + 1] + A.cs[id + sx + 1]);
      sig = A.sxz; rr = A.r[4];
    } else {                  // dvz/dy (D+), dvy/dz (D+)
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        d1 += dterm<FAST, 1>(A, c, A.vz[id + (c + 1) * sy], A.vz[id - c * sy]);
        d2 += dterm<FAST, 2>(A, c, A.vy[id + c + 1], A.vy[id - c]);
      }
      d1 = cpml_stretch_m<FAST>(A.ax[1].ah, A.ax[1].bh, A.ax[1].kh, A.ax[1].khi, A.psi[P_DVZ_DY], py, jy, by >= 0, d1);
      d2 = cpml_stretch_m<FAST>(A.ax[2].ah, A.ax[2].bh, A.ax[2].kh, A.ax[2].khi, A.psi[P_DVY_DZ], pz, jz, bz >= 0, d2);
      rho4 = 0.25f * (A.rho[id + sy] + A.rho[id] + A.rho[id + 1] + A.rho[id + sy + 1]);
      cs4 = 0.25f * (A.cs[id + sy] + A.cs[id] + A.cs[id + 1] + A.cs[id + sy + 1]);
      sig = A.syz; rr = A.r[5];
    }
    const float mu = rho4 * (cs4 * cs4);
    if constexpr (VISCO && LC > 0) {
      float o[LC];
#pragma unroll
      for (int l = 0; l < LC; l++) o[l] = rr[l * A.rs + id];
      const float sig0 = sig[id];
      float S = 0.f;
      const float X = d1 + d2;
#pragma unroll
      for (int l = 0; l < LC; l++) {
        const float n = (o[l] + X * A.zk.phi_s[l] - o[l] * A.zk.h_s[l]) * A.zk.m_s[l];
        S += o[l] + n;
        rr[l * A.rs + id] = n;
      }
      sig[id] = sig0 + (mu * (d1 + d2) + mu * 0.5f * S) * dt;
    } else if constexpr (VISCO) {
      float S = 0.f;
      const float X = d1 + d2;
      for (int l = 0; l < A.L; l++) {
        float *pr = rr + l * A.rs + id;
        const float o = *pr;
        const float n = (o + X * A.zk.phi_s[l] - o * A.zk.h_s[l]) * A.zk.m_s[l];
        S += o + n;
        *pr = n;
      }
      sig[id] = sig[id] + (mu * (d1 + d2) + mu * 0.5f * S) * dt;
    } else {
      sig[id] = sig[id] + (d2 + d1) * mu * dt;
    }
  }
}

template <int ORD, bool FAST>
__global__ __launch_bounds__(BZ * BYY) void k_velocity3d_plain(StepArgs A) {
  const int z = blockIdx.x * BZ + threadIdx.x;
  const int y = blockIdx.y * BYY + threadIdx.y;
  const int x = blockIdx.z;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  const bool valid = z < nz && y < ny;
  constexpr int last = ORD - 1, off = ORD - 1;
  const long sy = nz, sx = (long)ny * nz;
  const long id = x * sx + y * sy + z;
  const float dt = A.dt;
  unsigned v2bits = 0;
  if (valid) {
    const int jx = x - off, jy = y - off, jz = z - off;
    const int bx = band_index(A.ax[0], jx), by = band_index(A.ax[1], jy), bz = band_index(A.ax[2], jz);
    const long px = (long)bx * sx + y * sy + z;
    const long py = (long)x * A.ax[1].nb * nz + (long)by * nz + z;
    const long pz = ((long)x * ny + y) * A.ax[2].nb + bz;
    float vxn = A.vx[id], vyn = A.vy[id], vzn = A.vz[id];
    // vx (wgsl3d :1647-1681)
    if (x >= last + 1 && x < nx - last && y >= last + 1 && y < ny - last && z >= last + 1 && z < nz - last) {
      float a = 0.f, b = 0.f, c_ = 0.f;
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        a += dterm<FAST, 0>(A, c, A.sxx[id + c * sx], A.sxx[id - (c + 1) * sx]);
        b += dterm<FAST, 1>(A, c, A.sxy[id + c * sy], A.sxy[id - (c + 1) * sy]);
        c_ += dterm<FAST, 2>(A, c, A.sxz[id + c], A.sxz[id - c - 1]);
      }
      a = cpml_stretch_m<FAST>(A.ax[0].a, A.ax[0].b, A.ax[0].k, A.ax[0].ki, A.psi[P_DSXX_DX], px, jx, bx >= 0, a);
      b = cpml_stretch_m<FAST>(A.ax[1].a, A.ax[1].b, A.ax[1].k, A.ax[1].ki, A.psi[P_DSXY_DY], py, jy, by >= 0, b);
      c_ = cpml_stretch_m<FAST>(A.ax[2].a, A.ax[2].b, A.ax[2].k, A.ax[2].ki, A.psi[P_DSXZ_DZ], pz, jz, bz >= 0, c_);
      const float rho = 0.5f * (A.rho[id + sx] + A.rho[id]);
      if (rho > 0.f) vxn = (a + b + c_) * dt / rho + vxn;
    }
    // vy (wgsl3d :1683-1717)
    if (x >= last && x < nx - last - 1 && y >= last && y < ny - last - 1 && z >= last + 1 && z < nz - last) {
      float a = 0.f, b = 0.f, c_ = 0.f;
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        a += dterm<FAST, 0>(A, c, A.sxy[id + (c + 1) * sx], A.sxy[id - c * sx]);
        b += dterm<FAST, 1>(A, c, A.syy[id + (c + 1) * sy], A.syy[id - c * sy]);
        c_ += dterm<FAST, 2>(A, c, A.syz[id + c], A.syz[id - c - 1]);
      }
      a = cpml_stretch_m<FAST>(A.ax[0].ah, A.ax[0].bh, A.ax[0].kh, A.ax[0].khi, A.psi[P_DSXY_DX], px, jx, bx >= 0, a);
      b = cpml_stretch_m<FAST>(A.ax[1].ah, A.ax[1].bh, A.ax[1].kh, A.ax[1].khi, A.psi[P_DSYY_DY], py, jy, by >= 0, b);
      c_ = cpml_stretch_m<FAST>(A.ax[2].a, A.ax[2].b, A.ax[2].k, A.ax[2].ki, A.psi[P_DSYZ_DZ], pz, jz, bz >= 0, c_);
      const float rho = 0.5f * (A.rho[id + sy] + A.rho[id]);
      if (rho > 0.f) vyn = (a + b + c_) * dt / rho + vyn;
    }
    // vz, with the vy density (wgsl3d :1719-1758)
    if (x >= last && x < nx - last - 1 && y >= last + 1 && y < ny - last && z >= last && z < nz - last - 1) {
      float a = 0.f, b = 0.f, c_ = 0.f;
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        a += dterm<FAST, 0>(A, c, A.sxz[id + (c + 1) * sx], A.sxz[id - c * sx]);
        b += dterm<FAST, 1>(A, c, A.syz[id + c * sy], A.syz[id - (c + 1) * sy]);
        c_ += dterm<FAST, 2>(A, c, A.szz[id + c + 1], A.szz[id - c]);
      }
      a = cpml_stretch_m<FAST>(A.ax[0].ah, A.ax[0].bh, A.ax[0].kh, A.ax[0].khi, A.psi[P_DSXZ_DX], px, jx, bx >= 0, a);
      b = cpml_stretch_m<FAST>(A.ax[1].a, A.ax[1].b, A.ax[1].k, A.ax[1].ki, A.psi[P_DSYZ_DY], py, jy, by >= 0, b);
      c_ = cpml_stretch_m<FAST>(A.ax[2].ah, A.ax[2].bh, A.ax[2].kh, A.ax[2].khi, A.psi[P_DSZZ_DZ], pz, jz, bz >= 0, c_);
      const float rho = 0.5f * (A.rho[id + sy] + A.rho[id]);
      if (rho > 0.f) vzn = (a + b + c_) * dt / rho + vzn;
    } else {
      vxn = 0.f; vyn = 0.f; vzn = 0.f;
    }
    A.vx[id] = vxn;
    A.vy[id] = vyn;
    A.vz[id] = vzn;
    v2bits = __float_as_uint(vxn * vxn + vyn * vyn + vzn * vzn);
  }
  block_max_u32(v2bits, A.v2max);
}


// Launch shape: tiles of 64 (z) x TY (y) columns; x split into chunks of
// `xchunk` planes (1 = one plane per workgroup, i.e. one cell per thread).
struct Launch {
  int xchunk, nzb, nyb, nxc;
};

Launch launch_shape(const StepArgs &A, int ty, int xchunk) {
  Launch L;
  L.nzb = (A.nz + BZ - 1) / BZ;
  L.nyb = (A.ny + ty - 1) / ty;
  L.xchunk = std::max(1, std::min(xchunk, A.nx));
  L.nxc = (A.nx + L.xchunk - 1) / L.xchunk;
  return L;
}

template <int ORD, int TY, bool NT, bool FAST>
hipError_t launch_one(bool stress, int visco, const StepArgs &A, int xchunk, int remap, bool lds, hipStream_t st) {
  const Launch L = launch_shape(A, TY, xchunk);
  const dim3 grid((unsigned)(L.nzb * L.nyb * L.nxc)), block(BZ, TY);
  if (stress) {
    bool done = false;
    if constexpr (ORD == 2)
      if (visco && A.L == 3) {
        hipLaunchKernelGGL((k_stress3d<true, ORD, TY, NT, FAST, 3>), grid, block, 0, st, A, L.xchunk, L.nzb, L.nyb, L.nxc, remap);
        done = true;
      }
    if (done) {
    } else if (visco) hipLaunchKernelGGL((k_stress3d<true, ORD, TY, NT, FAST>), grid, block, 0, st, A, L.xchunk, L.nzb, L.nyb, L.nxc, remap);
    else hipLaunchKernelGGL((k_stress3d<false, ORD, TY, NT, FAST>), grid, block, 0, st, A, L.xchunk, L.nzb, L.nyb, L.nxc, remap);
  } else if (lds) {
    if constexpr (TY >= 2 * ORD)
      hipLaunchKernelGGL((k_velocity3d_lds<ORD, TY, NT, FAST>), grid, block, 0, st, A, L.xchunk, L.nzb, L.nyb, L.nxc, remap);
    else
      return hipErrorInvalidValue;
  } else {
    hipLaunchKernelGGL((k_velocity3d<ORD, TY, NT, FAST>), grid, block, 0, st, A, L.xchunk, L.nzb, L.nyb, L.nxc, remap);
  }
  return hipGetLastError();
}

template <int ORD, bool FAST>
hipError_t launch_tuned(bool stress, int visco, const StepArgs &A, const KTune &t, hipStream_t st) {
  // the x-march velocity kernel addresses whole fields with 32-bit offsets
  const bool plain = t.xchunk == 0 || (!stress && A.ncell >= (1L << 30));
  if (plain) {             // plain one-cell-per-thread kernels
    const dim3 block(BZ, BYY), grid((A.nz + BZ - 1) / BZ, (A.ny + BYY - 1) / BYY, A.nx);
    if (stress) {
      bool done = false;
      if constexpr (ORD == 2)
        if (visco && A.L == 3) {
          hipLaunchKernelGGL((k_stress3d_plain<true, ORD, FAST, 3>), grid, block, 0, st, A);
          done = true;
        }
      if (done) {
      } else if (visco) hipLaunchKernelGGL((k_stress3d_plain<true, ORD, FAST>), grid, block, 0, st, A);
      else hipLaunchKernelGGL((k_stress3d_plain<false, ORD, FAST>), grid, block, 0, st, A);
    } else {
      hipLaunchKernelGGL((k_velocity3d_plain<ORD, FAST>), grid, block, 0, st, A);
    }
    return hipGetLastError();
  }
  // compiled shapes: TY in {1, 4, 8} x non-temporal on/off (nt bit 0); nt bit 1
  // stages the velocity pass's y-neighbours in LDS (TY >= 2*ORD)
  const bool nt = (t.nt & 1) != 0, lds = !stress && (t.nt & 2) != 0;
  if (t.ty == 1) return nt ? launch_one<ORD, 1, true, FAST>(stress, visco, A, t.xchunk, t.remap, lds, st)
                           : launch_one<ORD, 1, false, FAST>(stress, visco, A, t.xchunk, t.remap, lds, st);
  if (t.ty == 4) return nt ? launch_one<ORD, 4, true, FAST>(stress, visco, A, t.xchunk, t.remap, lds, st)
                           : launch_one<ORD, 4, false, FAST>(stress, visco, A, t.xchunk, t.remap, lds, st);
  if (t.ty == 8) return nt ? launch_one<ORD, 8, true, FAST>(stress, visco, A, t.xchunk, t.remap, lds, st)
                           : launch_one<ORD, 8, false, FAST>(stress, visco, A, t.xchunk, t.remap, lds, st);
  return hipErrorInvalidValue;
}

template <bool FAST>
hipError_t dispatch(bool stress, int visco, int ord, const StepArgs &A, const KTune &t, hipStream_t st) {
  switch (ord) {
    case 2: return launch_tuned<2, FAST>(stress, visco, A, t, st);
    case 3: return launch_tuned<3, FAST>(stress, visco, A, t, st);
    case 4: return launch_tuned<4, FAST>(stress, visco, A, t, st);
    case 5: return launch_tuned<5, FAST>(stress, visco, A, t, st);
    case 6: return launch_tuned<6, FAST>(stress, visco, A, t, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

hipError_t vse_launch_stress3d(int visco, int ord, int math, const StepArgs &A, const KTune &t, hipStream_t st) {
  return math == VSE_MATH_FAST ? dispatch<true>(true, visco, ord, A, t, st) : dispatch<false>(true, visco, ord, A, t, st);
}

hipError_t vse_launch_velocity3d(int visco, int ord, int math, const StepArgs &A, const KTune &t, hipStream_t st) {
  return math == VSE_MATH_FAST ? dispatch<true>(false, visco, ord, A, t, st) : dispatch<false>(false, visco, ord, A, t, st);
}
