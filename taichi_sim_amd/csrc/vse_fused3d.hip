// vse_fused3d.hip -- fused 3D step: the velocity update of step n and the
// stress update of step n+1 in one x-march (viscoelastic, 3 mechanisms, FD
// order 2).
//
// The two-pass step (vse_k3d.hip) moves sigma and v through HBM twice per
// step: the stress pass reads v and rewrites sigma, the velocity pass reads
// sigma and rewrites v.  Here one kernel does both for each x-plane of a
// tile, so each field crosses HBM once per step.
//
// * A workgroup owns a 64 (z) x TY (y) tile and marches x over a chunk.  At
//   iteration x it stages plane xv = x + 2 of sigma(n+1) over the tile plus 4
//   cells on each side in LDS, computes v(n+1) of plane xv over the tile plus
//   2 cells on each side (the halo the stress stencils need) into a 3-plane
//   LDS ring, then updates the stress of plane x from v(n+1) of planes
//   x-2 .. x+2.  The march starts 4 planes early to fill the ring and the
//   x-windows.
// * Halo velocities belong to other workgroups, which compute them too: every
//   copy comes from the same inputs through the same arithmetic, so all have
//   the same bits.  The inputs must not change meanwhile, so sigma, v and the
//   velocity-pass psi are double-buffered: read from one set, written to the
//   other.  Memory variables and the stress-pass psi are touched only at the
//   own cell and stay in place.
// * Every expression is the two-pass kernels' (k_velocity3d, then k_stress3d
//   with 3 mechanisms; shader_webgpu_3D_viscoelastic.wgsl:1633-1773 and
//   :1340-1631), in the same order, so the fused step is bit-identical to the
//   two-pass step.
// * WGSL-variant sources go into vz right after its update (wgsl3d
//   :1760-1766; the two-pass step does it in the sparse kernel between the
//   passes), found through a per-(x, y)-row table.
#include "vse_device.h"

namespace {

constexpr int ORD = 2;            // FD coefficients
constexpr int H = ORD;            // velocity halo the stress stencils read
constexpr int LC = 3;             // Zener mechanisms
constexpr int BZ = 64;
constexpr int EZ = BZ + 2 * H;    // E-region (velocities: tile + halo), z extent
constexpr int SZ = BZ + 4 * H;    // S-region (staged sigma: E-region + stencil), z extent

// A whole field as a raw buffer: range-checked, 32-bit element offsets (the
// host launches this kernel only when every array has fewer than 2^32 bytes).
struct RB {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ RB(const float *base, unsigned nbytes)
      : r(__builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)nbytes, 0x00020000)) {}
  __device__ __forceinline__ float ld(unsigned i) const {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)(i * 4u), 0, 0));
  }
  __device__ __forceinline__ void st(unsigned i, float v) const {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)(i * 4u), 0, 0);
  }
};

// One staggered-difference term along axis AX, as in vse_k3d.hip.
template <bool FAST, int AX>
__device__ __forceinline__ float dterm(const StepArgs &A, int c, float a, float b) {
  if constexpr (FAST) return (a - b) * A.Cd[AX][c];
  else return A.C[c] * (a - b) / (AX == 0 ? A.dx : (AX == 1 ? A.dy : A.dz));
}

// psi <- b psi + a D (read from `in`, written to `out` when `store`), returns
// D/k + psi; FAST multiplies by the precomputed 1/k (wgsl3d :1372-1378).
template <bool FAST>
__device__ __forceinline__ float stretch(const float *a, const float *b, const float *k, const float *ki,
                                         const float *in, float *out, unsigned i, int j, float D, bool store) {
  const float pn = b[j] * in[i] + a[j] * D;
  if (store) out[i] = pn;
  if constexpr (FAST) return D * ki[j] + pn;
  else return D / k[j] + pn;
}

__device__ __forceinline__ int mod3(int p) { return ((p % 3) + 3) % 3; }

template <bool FAST, int TY>
__global__ __launch_bounds__(BZ * TY) void k_fused3d(FusedArgs F, int xchunk, int nzb, int nyb, int nxc, int remap) {
  constexpr int NT = BZ * TY;
  constexpr int EY = TY + 2 * H, NE = EY * EZ;
  constexpr int SY = TY + 4 * H, NS = SY * SZ;
  constexpr int RND = (NE + NT - 1) / NT;          // E-cells per thread
  constexpr int last = ORD - 1, off = ORD - 1;
  // plane xv of sigma(n+1) xy, xz, yy, yz, zz over the S-region; cell (ys, zs)
  // is grid (y0 - 2H + ys, z0 - 2H + zs), out-of-grid cells 0
  __shared__ float sS[5][NS];
  // v(n+1) of plane p in slot p mod 3, components x, y, z, over the E-region;
  // cell (ey, ez) is grid (y0 - H + ey, z0 - H + ez), out-of-grid cells 0
  __shared__ float sV[3][3][NE];

  const StepArgs &A = F.A;
  const Tile T = tile_of_block(nzb, nyb, nxc, remap);
  const int tid = threadIdx.x + threadIdx.y * BZ;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  const int z0 = T.bz * BZ, y0 = T.by * TY;
  const int x0 = T.bxc * xchunk, x1 = min(nx, x0 + xchunk);
  const unsigned sy = nz, sx = (unsigned)ny * nz;
  const unsigned fb = (unsigned)(A.ncell * 4);
  const float dt = A.dt;
  const int s = *A.step;                            // the step the velocity part advances
  const int nby = A.ax[1].nb, nbz = A.ax[2].nb;
  // the streamed inputs through buffer descriptors; outputs, memory variables
  // and psi through plain pointers (fewer scalar registers live in the loop)
  const RB Isxx(A.sxx, fb), Isyy(A.syy, fb), Iszz(A.szz, fb), Isxy(A.sxy, fb), Isxz(A.sxz, fb), Isyz(A.syz, fb);
  const RB Ivx(A.vx, fb), Ivy(A.vy, fb), Ivz(A.vz, fb), Brho(A.rho, fb), Bcp(A.cp, fb), Bcs(A.cs, fb);
  float *const Ovx = F.vo[0], *const Ovy = F.vo[1], *const Ovz = F.vo[2];
  // first element of plane x, clamped into the grid (clamped planes feed no update)
  auto pl = [&](int x) { return (unsigned)(x < 0 ? 0 : (x >= nx ? nx - 1 : x)) * sx; };

  // x-windows of sigma at this thread's E-cells (c = tid + k*NT): sigma_xx at
  // planes xv-ORD .. xv+ORD-1 (D-), sigma_xy / sigma_xz at xv-ORD+1 .. xv+ORD (D+)
  float wxx[RND][2 * ORD], wxy[RND][2 * ORD], wxz[RND][2 * ORD];
  // v(n+1) of the own column at planes x-ORD .. x+ORD (the stress x-windows)
  float svx[2 * ORD + 1], svy[2 * ORD + 1], svz[2 * ORD + 1];
#pragma unroll
  for (int i = 0; i < 2 * ORD + 1; i++) svx[i] = svy[i] = svz[i] = 0.f;
  unsigned v2bits = 0;

  {
    const int xs = x0 - H;                          // first velocity plane
#pragma unroll
    for (int k = 0; k < RND; k++) {
      const int c = tid + k * NT;
      const int ey = c / EZ, ez = c - (c / EZ) * EZ;
      const int y = y0 - H + ey, z = z0 - H + ez;
      const bool in = c < NE && y >= 0 && y < ny && z >= 0 && z < nz;
      const unsigned col = in ? (unsigned)y * sy + (unsigned)z : 0u;
#pragma unroll
      for (int i = 0; i < 2 * ORD - 1; i++) {
        wxx[k][i + 1] = in ? Isxx.ld(pl(xs - ORD + i) + col) : 0.f;
        wxy[k][i + 1] = in ? Isxy.ld(pl(xs - (ORD - 1) + i) + col) : 0.f;
        wxz[k][i + 1] = in ? Isxz.ld(pl(xs - (ORD - 1) + i) + col) : 0.f;
      }
    }
  }

#pragma unroll 1
  for (int x = x0 - 2 * H; x < x1; x++) {
    const int xv = x + H;
    const int vslot = mod3(xv);
    const bool xv_in = xv >= 0 && xv < nx;
    // ---- stage sigma(n+1) of plane xv over the S-region
    {
      const unsigned pb = xv_in ? (unsigned)xv * sx : 0u;
      for (int i = tid; i < NS; i += NT) {
        const int ys = i / SZ, zs = i - (i / SZ) * SZ;
        const int y = y0 - 2 * H + ys, z = z0 - 2 * H + zs;
        const bool in = xv_in && y >= 0 && y < ny && z >= 0 && z < nz;
        const unsigned g = pb + (in ? (unsigned)y * sy + (unsigned)z : 0u);
        float a = 0.f, b = 0.f, c = 0.f, d = 0.f, e = 0.f;
        if (in) {
          a = Isxy.ld(g); b = Isxz.ld(g); c = Isyy.ld(g); d = Isyz.ld(g); e = Iszz.ld(g);
        }
        sS[0][i] = a; sS[1][i] = b; sS[2][i] = c; sS[3][i] = d; sS[4][i] = e;
      }
    }
    __syncthreads();

    // ---- v(n+1) of plane xv over the E-region (k_velocity3d's arithmetic)
#pragma unroll
    for (int k = 0; k < RND; k++) {
      const int c = tid + k * NT;
      if (c >= NE) continue;
      const int ey = c / EZ, ez = c - (c / EZ) * EZ;
      const int y = y0 - H + ey, z = z0 - H + ez;
      const bool in = y >= 0 && y < ny && z >= 0 && z < nz;
      const unsigned col = in ? (unsigned)y * sy + (unsigned)z : 0u;
#pragma unroll
      for (int i = 0; i < 2 * ORD - 1; i++) {
        wxx[k][i] = wxx[k][i + 1]; wxy[k][i] = wxy[k][i + 1]; wxz[k][i] = wxz[k][i + 1];
      }
      wxx[k][2 * ORD - 1] = in ? Isxx.ld(pl(xv + ORD - 1) + col) : 0.f;
      wxy[k][2 * ORD - 1] = in ? Isxy.ld(pl(xv + ORD) + col) : 0.f;
      wxz[k][2 * ORD - 1] = in ? Isxz.ld(pl(xv + ORD) + col) : 0.f;
      float vxn = 0.f, vyn = 0.f, vzn = 0.f;
      if (in && xv_in) {
        const bool own = xv >= x0 && xv < x1 && ey >= H && ey < H + TY && ez >= H && ez < H + BZ;
        const unsigned i0 = (unsigned)xv * sx + col;
        vxn = Ivx.ld(i0); vyn = Ivy.ld(i0); vzn = Ivz.ld(i0);
        const float rho0 = Brho.ld(i0), rho_n = Brho.ld(pl(xv + 1) + col), rho_y1 = Brho.ld(i0 + sy);
        const int is = (ey + H) * SZ + (ez + H);    // the E-cell in the S-region
        const int jx = xv - off, jy = y - off, jz = z - off;
        const int bx = band_index(A.ax[0], jx), by = band_index(A.ax[1], jy), bz = band_index(A.ax[2], jz);
        const unsigned pxi = (unsigned)bx * sx + col;
        const unsigned pyi = (unsigned)xv * nby * nz + (unsigned)by * nz + (unsigned)z;
        const unsigned pzi = ((unsigned)xv * ny + (unsigned)y) * nbz + (unsigned)bz;
        const bool yi_vx = y >= last + 1 && y < ny - last, zi_vx = z >= last + 1 && z < nz - last;
        const bool yi_vy = y >= last && y < ny - last - 1, zi_vy = z >= last + 1 && z < nz - last;
        const bool yi_vz = y >= last + 1 && y < ny - last, zi_vz = z >= last && z < nz - last - 1;
        // vx (wgsl3d :1647-1681)
        if (xv >= last + 1 && xv < nx - last && yi_vx && zi_vx) {
          float a = 0.f, b = 0.f, c_ = 0.f;
#pragma unroll
          for (int q = 0; q < ORD; q++) {
            a += dterm<FAST, 0>(A, q, wxx[k][ORD + q], wxx[k][ORD - 1 - q]);       // sxx[x+q] - sxx[x-q-1]
            b += dterm<FAST, 1>(A, q, sS[0][is + q * SZ], sS[0][is - (q + 1) * SZ]);
            c_ += dterm<FAST, 2>(A, q, sS[1][is + q], sS[1][is - q - 1]);
          }
          if (bx >= 0)
            a = stretch<FAST>(A.ax[0].a, A.ax[0].b, A.ax[0].k, A.ax[0].ki, A.psi[P_DSXX_DX],
                              F.psio[P_DSXX_DX - P_DSXX_DX], pxi, jx, a, own);
          if (by >= 0)
            b = stretch<FAST>(A.ax[1].a, A.ax[1].b, A.ax[1].k, A.ax[1].ki, A.psi[P_DSXY_DY],
                              F.psio[P_DSXY_DY - P_DSXX_DX], pyi, jy, b, own);
          if (bz >= 0)
            c_ = stretch<FAST>(A.ax[2].a, A.ax[2].b, A.ax[2].k, A.ax[2].ki, A.psi[P_DSXZ_DZ],
                               F.psio[P_DSXZ_DZ - P_DSXX_DX], pzi, jz, c_, own);
          const float rho = 0.5f * (rho_n + rho0);
          if (rho > 0.f) vxn = (a + b + c_) * dt / rho + vxn;
        }
        // vy (wgsl3d :1683-1717)
        const bool xi_vyz = xv >= last && xv < nx - last - 1;
        if (xi_vyz && yi_vy && zi_vy) {
          float a = 0.f, b = 0.f, c_ = 0.f;
#pragma unroll
          for (int q = 0; q < ORD; q++) {
            a += dterm<FAST, 0>(A, q, wxy[k][ORD + q], wxy[k][ORD - 1 - q]);       // sxy[x+q+1] - sxy[x-q]
            b += dterm<FAST, 1>(A, q, sS[2][is + (q + 1) * SZ], sS[2][is - q * SZ]);
            c_ += dterm<FAST, 2>(A, q, sS[3][is + q], sS[3][is - q - 1]);
          }
          if (bx >= 0)
            a = stretch<FAST>(A.ax[0].ah, A.ax[0].bh, A.ax[0].kh, A.ax[0].khi, A.psi[P_DSXY_DX],
                              F.psio[P_DSXY_DX - P_DSXX_DX], pxi, jx, a, own);
          if (by >= 0)
            b = stretch<FAST>(A.ax[1].ah, A.ax[1].bh, A.ax[1].kh, A.ax[1].khi, A.psi[P_DSYY_DY],
                              F.psio[P_DSYY_DY - P_DSXX_DX], pyi, jy, b, own);
          if (bz >= 0)
            c_ = stretch<FAST>(A.ax[2].a, A.ax[2].b, A.ax[2].k, A.ax[2].ki, A.psi[P_DSYZ_DZ],
                               F.psio[P_DSYZ_DZ - P_DSXX_DX], pzi, jz, c_, own);
          const float rho = 0.5f * (rho_y1 + rho0);
          if (rho > 0.f) vyn = (a + b + c_) * dt / rho + vyn;
        }
        // vz, with the vy density (wgsl3d :1719-1758)
        if (xi_vyz && yi_vz && zi_vz) {
          float a = 0.f, b = 0.f, c_ = 0.f;
#pragma unroll
          for (int q = 0; q < ORD; q++) {
            a += dterm<FAST, 0>(A, q, wxz[k][ORD + q], wxz[k][ORD - 1 - q]);       // sxz[x+q+1] - sxz[x-q]
            b += dterm<FAST, 1>(A, q, sS[3][is + q * SZ], sS[3][is - (q + 1) * SZ]);
            c_ += dterm<FAST, 2>(A, q, sS[4][is + q + 1], sS[4][is - q]);
          }
          if (bx >= 0)
            a = stretch<FAST>(A.ax[0].ah, A.ax[0].bh, A.ax[0].kh, A.ax[0].khi, A.psi[P_DSXZ_DX],
                              F.psio[P_DSXZ_DX - P_DSXX_DX], pxi, jx, a, own);
          if (by >= 0)
            b = stretch<FAST>(A.ax[1].a, A.ax[1].b, A.ax[1].k, A.ax[1].ki, A.psi[P_DSYZ_DY],
                              F.psio[P_DSYZ_DY - P_DSXX_DX], pyi, jy, b, own);
          if (bz >= 0)
            c_ = stretch<FAST>(A.ax[2].ah, A.ax[2].bh, A.ax[2].kh, A.ax[2].khi, A.psi[P_DSZZ_DZ],
                               F.psio[P_DSZZ_DZ - P_DSXX_DX], pzi, jz, c_, own);
          const float rho = 0.5f * (rho_y1 + rho0);
          if (rho > 0.f) vzn = (a + b + c_) * dt / rho + vzn;
        } else {
          vxn = 0.f; vyn = 0.f; vzn = 0.f;
        }
        if (own) {
          const unsigned u = __float_as_uint(vxn * vxn + vyn * vyn + vzn * vzn);
          v2bits = v2bits > u ? v2bits : u;
        }
        // source into vz with the vy-grid density, out-of-grid neighbour 0 (wgsl3d :1760-1766)
        if (A.src_rbeg && s < A.src_rows && A.src_rowflag[xv]) {
          const int row = xv * ny + y;
          const int q1 = A.src_rbeg[row + 1];
          for (int q = A.src_rbeg[row]; q < q1; q++) {
            if (A.src_rz[q] != z) continue;
            const float rho = 0.5f * ((y + 1 < ny ? rho_y1 : 0.f) + rho0);
            if (rho > 0.f) {
              const float val = A.src[(long)s * A.src_cols + A.src_rcol[q]];
              vzn = vzn + val * dt / rho;
              float m = vxn * vxn + vyn * vyn;
              m = m + vzn * vzn;
              const unsigned u = __float_as_uint(m);
              if (own) v2bits = v2bits > u ? v2bits : u;
            }
          }
        }
        if (own) {
          Ovx[i0] = vxn; Ovy[i0] = vyn; Ovz[i0] = vzn;
        }
      }
      sV[vslot][0][c] = vxn; sV[vslot][1][c] = vyn; sV[vslot][2][c] = vzn;
    }
    __syncthreads();

    // the own column's v(n+1) of plane xv enters the stress x-windows
    const int ie = (threadIdx.y + H) * EZ + threadIdx.x + H;
#pragma unroll
    for (int i = 0; i < 2 * ORD; i++) { svx[i] = svx[i + 1]; svy[i] = svy[i + 1]; svz[i] = svz[i + 1]; }
    svx[2 * ORD] = sV[vslot][0][ie]; svy[2 * ORD] = sV[vslot][1][ie]; svz[2 * ORD] = sV[vslot][2][ie];

    // ---- stress of plane x from v(n+1) (k_stress3d's arithmetic, 3 mechanisms)
    const int yo = y0 + threadIdx.y, zo = z0 + threadIdx.x;
    if (x >= x0 && yo < ny && zo < nz) {
      const int cs_ = mod3(x);
      const float *V0 = sV[cs_][0], *V1 = sV[cs_][1], *V2 = sV[cs_][2];
      const unsigned col = (unsigned)yo * sy + (unsigned)zo, i0 = (unsigned)x * sx + col, i1 = pl(x + 1) + col;
      const int jx = x - off, jy = yo - off, jz = zo - off;
      const int bx = band_index(A.ax[0], jx), by = band_index(A.ax[1], jy), bz = band_index(A.ax[2], jz);
      const unsigned pxi = (unsigned)bx * sx + col;
      const unsigned pyi = (unsigned)x * nby * nz + (unsigned)by * nz + (unsigned)zo;
      const unsigned pzi = ((unsigned)x * ny + (unsigned)yo) * nbz + (unsigned)bz;
      const bool xp = x >= last && x < nx - last - 1;
      const bool yp = yo >= last && yo < ny - last - 1;
      const bool zp = zo >= last && zo < nz - last - 1;
      const bool xm = x >= last + 1 && x < nx - last;
      const bool ym = yo >= last + 1 && yo < ny - last;
      const bool zm = zo >= last + 1 && zo < nz - last;
      float nvy_p[ORD], nvy_m[ORD], nvz_p[ORD], nvz_m[ORD];        // normal: dvy (D-), dvz (D-)
      float nxy_p[ORD], nxy_m[ORD], nxz_p[ORD], nxz_m[ORD];        // shear: dvx/dy, dvx/dz (D+)
      float nzy_p[ORD], nzy_m[ORD], nyz_p[ORD], nyz_m[ORD];        // shear: dvz/dy, dvy/dz (D+)
#pragma unroll
      for (int q = 0; q < ORD; q++) {
        nvy_p[q] = V1[ie + q * EZ]; nvy_m[q] = V1[ie - (q + 1) * EZ];
        nvz_p[q] = V2[ie + q]; nvz_m[q] = V2[ie - q - 1];
        nxy_p[q] = V0[ie + (q + 1) * EZ]; nxy_m[q] = V0[ie - q * EZ];
        nxz_p[q] = V0[ie + q + 1]; nxz_m[q] = V0[ie - q];
        nzy_p[q] = V2[ie + (q + 1) * EZ]; nzy_m[q] = V2[ie - q * EZ];
        nyz_p[q] = V1[ie + q + 1]; nyz_m[q] = V1[ie - q];
      }
      const float rho0 = Brho.ld(i0), cs0 = Bcs.ld(i0), rho_n = Brho.ld(i1), cs_n = Bcs.ld(i1);
      const float m_r0y = Brho.ld(i0 + sy), m_r1y = Brho.ld(i1 + sy), m_c0y = Bcs.ld(i0 + sy), m_c1y = Bcs.ld(i1 + sy);
      const float m_r0z = Brho.ld(i0 + 1), m_r1z = Brho.ld(i1 + 1), m_c0z = Bcs.ld(i0 + 1), m_c1z = Bcs.ld(i1 + 1);
      const float m_r0yz = Brho.ld(i0 + sy + 1), m_c0yz = Bcs.ld(i0 + sy + 1), m_cp = Bcp.ld(i0);
      float *const Osxx = F.so[0], *const Osyy = F.so[1], *const Oszz = F.so[2];

      // ---- normal stresses (wgsl3d :1356-1453): x D+, y D-, z D-
      {
        const float sxx0 = Isxx.ld(i0), syy0 = Isyy.ld(i0), szz0 = Iszz.ld(i0);
        if (xp && ym && zm) {
          float dvx = 0.f, dvy = 0.f, dvz = 0.f;
#pragma unroll
          for (int q = 0; q < ORD; q++) {
            dvx += dterm<FAST, 0>(A, q, svx[1 + ORD + q], svx[ORD - q]);           // vx[x+q+1] - vx[x-q]
            dvy += dterm<FAST, 1>(A, q, nvy_p[q], nvy_m[q]);
            dvz += dterm<FAST, 2>(A, q, nvz_p[q], nvz_m[q]);
          }
          if (bx >= 0) {
            float *P = A.psi[P_DVX_DX];
            dvx = stretch<FAST>(A.ax[0].ah, A.ax[0].bh, A.ax[0].kh, A.ax[0].khi, P, P, pxi, jx, dvx, true);
          }
          if (by >= 0) {
            float *P = A.psi[P_DVY_DY];
            dvy = stretch<FAST>(A.ax[1].a, A.ax[1].b, A.ax[1].k, A.ax[1].ki, P, P, pyi, jy, dvy, true);
          }
          if (bz >= 0) {
            float *P = A.psi[P_DVZ_DZ];
            dvz = stretch<FAST>(A.ax[2].a, A.ax[2].b, A.ax[2].k, A.ax[2].ki, P, P, pzi, jz, dvz, true);
          }
          const float cp = m_cp;
          const float lambda = rho0 * (cp * cp - 2.0f * cs0 * cs0);
          const float mu = rho0 * (cs0 * cs0);
          const float lp2m = lambda + 2.0f * mu;
          float *const R0 = A.r[0], *const R1 = A.r[1], *const R2 = A.r[2];
          float oxx[LC], oyy[LC], ozz[LC];
#pragma unroll
          for (int l = 0; l < LC; l++) {
            const unsigned li = (unsigned)(l * A.rs) + i0;
            oxx[l] = R0[li]; oyy[l] = R1[li]; ozz[l] = R2[li];
          }
          float Sxx = 0.f, Syy = 0.f, Szz = 0.f;
          const float div = dvx + dvy + dvz;
          const float div3 = FAST ? div * (1.0f / 3.0f) : div / 3.0f;
#pragma unroll
          for (int l = 0; l < LC; l++) {
            const float nxx = ((oxx[l] + div * A.zk.phi_p[l] - oxx[l] * A.zk.h_p[l]) * A.zk.m_p[l]);
            const float nyy = ((oyy[l] + (dvx - div3) * A.zk.phi_s[l] - oyy[l] * A.zk.h_s[l]) * A.zk.m_s[l]);
            const float nzz = ((ozz[l] + (dvy - div3) * A.zk.phi_s[l] - ozz[l] * A.zk.h_s[l]) * A.zk.m_s[l]);
            Sxx += oxx[l] + nxx; Syy += oyy[l] + nyy; Szz += ozz[l] + nzz;
            const unsigned li = (unsigned)(l * A.rs) + i0;
            R0[li] = nxx; R1[li] = nyy; R2[li] = nzz;
          }
          const float l23 = lambda + 2.0f * mu / 3.0f;
          Osxx[i0] = sxx0 + (lp2m * dvx + lambda * (dvy + dvz) + l23 * 0.5f * Sxx + mu * Syy) * dt;
          Osyy[i0] = syy0 + (lambda * (dvx + dvz) + lp2m * dvy + l23 * 0.5f * Sxx + mu * Szz) * dt;
          Oszz[i0] = szz0 + (lambda * (dvx + dvy) + lp2m * dvz + lp2m * 0.5f * Sxx - (mu / 3.0f) * (Syy + Szz)) * dt;
        } else {
          Osxx[i0] = sxx0; Osyy[i0] = syy0; Oszz[i0] = szz0;
        }
      }

      // ---- shear stresses (wgsl3d :1457-1629)
#pragma unroll
      for (int sh = 0; sh < 3; sh++) {
        const RB Isig(sh == 0 ? A.sxy : (sh == 1 ? A.sxz : A.syz), fb);
        float *const Osig = F.so[3 + sh];
        const float sig0 = Isig.ld(i0);
        const bool ok = (sh < 2 ? xm : xp) && yp && zp;
        if (!ok) {
          Osig[i0] = sig0;
          continue;
        }
        float d1 = 0.f, d2 = 0.f, rho4, cs4;
        const VSE_G float *rr;
        if (sh == 0) {            // sigma_xy: dvy/dx (D-), dvx/dy (D+)
#pragma unroll
          for (int q = 0; q < ORD; q++) {
            d1 += dterm<FAST, 0>(A, q, svy[ORD + q], svy[ORD - 1 - q]);           // vy[x+q] - vy[x-q-1]
            d2 += dterm<FAST, 1>(A, q, nxy_p[q], nxy_m[q]);
          }
          if (bx >= 0) {
            float *P = A.psi[P_DVY_DX];
            d1 = stretch<FAST>(A.ax[0].a, A.ax[0].b, A.ax[0].k, A.ax[0].ki, P, P, pxi, jx, d1, true);
          }
          if (by >= 0) {
            float *P = A.psi[P_DVX_DY];
            d2 = stretch<FAST>(A.ax[1].ah, A.ax[1].bh, A.ax[1].kh, A.ax[1].khi, P, P, pyi, jy, d2, true);
          }
          rho4 = 0.25f * (rho_n + rho0 + m_r0y + m_r1y);
          cs4 = 0.25f * (cs_n + cs0 + m_c0y + m_c1y);
          rr = A.r[3];
        } else if (sh == 1) {     // sigma_xz: dvz/dx (D-), dvx/dz (D+)
#pragma unroll
          for (int q = 0; q < ORD; q++) {
            d1 += dterm<FAST, 0>(A, q, svz[ORD + q], svz[ORD - 1 - q]);           // vz[x+q] - vz[x-q-1]
            d2 += dterm<FAST, 2>(A, q, nxz_p[q], nxz_m[q]);
          }
          if (bx >= 0) {
            float *P = A.psi[P_DVZ_DX];
            d1 = stretch<FAST>(A.ax[0].a, A.ax[0].b, A.ax[0].k, A.ax[0].ki, P, P, pxi, jx, d1, true);
          }
          if (bz >= 0) {
            float *P = A.psi[P_DVX_DZ];
            d2 = stretch<FAST>(A.ax[2].ah, A.ax[2].bh, A.ax[2].kh, A.ax[2].khi, P, P, pzi, jz, d2, true);
          }
          rho4 = 0.25f * (rho_n + rho0 + m_r0z + m_r1z);
          cs4 = 0.25f * (cs_n + cs0 + m_c0z + m_c1z);
          rr = A.r[4];
        } else {                  // sigma_yz: dvz/dy (D+), dvy/dz (D+)
#pragma unroll
          for (int q = 0; q < ORD; q++) {
            d1 += dterm<FAST, 1>(A, q, nzy_p[q], nzy_m[q]);
            d2 += dterm<FAST, 2>(A, q, nyz_p[q], nyz_m[q]);
          }
          if (by >= 0) {
            float *P = A.psi[P_DVZ_DY];
            d1 = stretch<FAST>(A.ax[1].ah, A.ax[1].bh, A.ax[1].kh, A.ax[1].khi, P, P, pyi, jy, d1, true);
          }
          if (bz >= 0) {
            float *P = A.psi[P_DVY_DZ];
            d2 = stretch<FAST>(A.ax[2].ah, A.ax[2].bh, A.ax[2].kh, A.ax[2].khi, P, P, pzi, jz, d2, true);
          }
          rho4 = 0.25f * (m_r0y + rho0 + m_r0z + m_r0yz);
          cs4 = 0.25f * (m_c0y + cs0 + m_c0z + m_c0yz);
          rr = A.r[5];
        }
        const float mu = rho4 * (cs4 * cs4);
        float *const R = (float *)rr;
        float o[LC];
#pragma unroll
        for (int l = 0; l < LC; l++) o[l] = R[(unsigned)(l * A.rs) + i0];
        float S = 0.f;
        const float X = d1 + d2;
#pragma unroll
        for (int l = 0; l < LC; l++) {
          const float n = (o[l] + X * A.zk.phi_s[l] - o[l] * A.zk.h_s[l]) * A.zk.m_s[l];
          S += o[l] + n;
          R[(unsigned)(l * A.rs) + i0] = n;
        }
        Osig[i0] = sig0 + (mu * (d1 + d2) + mu * 0.5f * S) * dt;
      }
    }
  }
  block_max_u32(v2bits, A.v2max);
}

}  // namespace

bool vse_fused3d_supported(long ncell, long rs, int L, int visco, int ord) {
  return visco && L == LC && ord == ORD && ncell > 0 && rs >= ncell && rs * 4L * LC < (1L << 32);
}

hipError_t vse_launch_fused3d(int math, const FusedArgs &F, const KTune &t, hipStream_t st) {
  const StepArgs &A = F.A;
  if (!vse_fused3d_supported(A.ncell, A.rs, A.L, 1, ORD)) return hipErrorInvalidValue;
  const int ty = t.ty == 8 ? 8 : 4;
  const int nzb = (A.nz + BZ - 1) / BZ, nyb = (A.ny + ty - 1) / ty;
  const int xchunk = std::max(1, std::min(t.xchunk > 0 ? t.xchunk : 16, A.nx));
  const int nxc = (A.nx + xchunk - 1) / xchunk;
  const dim3 grid((unsigned)(nzb * nyb * nxc)), block(BZ, ty);
  const bool fast = math == VSE_MATH_FAST;
  if (ty == 8) {
    if (fast) hipLaunchKernelGGL((k_fused3d<true, 8>), grid, block, 0, st, F, xchunk, nzb, nyb, nxc, t.remap);
    else hipLaunchKernelGGL((k_fused3d<false, 8>), grid, block, 0, st, F, xchunk, nzb, nyb, nxc, t.remap);
  } else {
    if (fast) hipLaunchKernelGGL((k_fused3d<true, 4>), grid, block, 0, st, F, xchunk, nzb, nyb, nxc, t.remap);
    else hipLaunchKernelGGL((k_fused3d<false, 4>), grid, block, 0, st, F, xchunk, nzb, nyb, nxc, t.remap);
  }
  return hipGetLastError();
}
