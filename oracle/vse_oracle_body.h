/*
 * vse_oracle_body.h -- included twice by vse_oracle.c with REAL = float and
 * REAL = double.  TEST INFRASTRUCTURE (see vse_oracle.h).
 *
 * Literal restatement of the reference kernels, one cell at a time:
 * accessors return 0 outside the grid (shader_webgpu_viscoelastic.wgsl:18-22,
 * :336-340; shader_webgpu_3D_viscoelastic.wgsl:28-32), CPML profile reads
 * return 0 outside [0, n - 2*(ord-1)) (wgsl2d :92-96), and every expression
 * keeps the reference's operand order so FP32 rounding matches term by term.
 */

#define K(c) ((REAL)(c))

typedef struct {
    REAL *vx, *vy, *vz;
    REAL *sxx, *syy, *szz, *sxy, *sxz, *syz;
    REAL *psi[PSI_COUNT];        /* full-grid memory variables (NULL if unused) */
    REAL *r[6];                  /* r_xx, r_yy, r_zz, r_xy, r_xz, r_yz: (cells, L) */
    REAL *traces;                /* (n_iter, n_rec_el) */
    REAL *v2max;                 /* running max |v|^2 (one scalar) */
} SFX(OracleState);

/* ---- accessors --------------------------------------------------------- */
static inline REAL SFX(g2)(const OracleParams *p, const REAL *f, int x, int y)
{
    if (x < 0 || x >= p->nx || y < 0 || y >= p->ny) return K(0.0);
    return f[(long)y + (long)x * p->ny];
}
static inline REAL SFX(m2)(const OracleParams *p, const float *f, int x, int y)
{
    if (x < 0 || x >= p->nx || y < 0 || y >= p->ny) return K(0.0);
    return (REAL)f[(long)y + (long)x * p->ny];
}
static inline long SFX(i3)(const OracleParams *p, int x, int y, int z)
{
    return (long)z + (long)y * p->nz + (long)x * p->nz * p->ny;
}
static inline int SFX(in3)(const OracleParams *p, int x, int y, int z)
{
    return x >= 0 && x < p->nx && y >= 0 && y < p->ny && z >= 0 && z < p->nz;
}
static inline REAL SFX(g3)(const OracleParams *p, const REAL *f, int x, int y, int z)
{
    if (!SFX(in3)(p, x, y, z)) return K(0.0);
    return f[SFX(i3)(p, x, y, z)];
}
static inline REAL SFX(m3)(const OracleParams *p, const float *f, int x, int y, int z)
{
    if (!SFX(in3)(p, x, y, z)) return K(0.0);
    return (REAL)f[SFX(i3)(p, x, y, z)];
}
/* CPML profile of one axis (0=x,1=y,2=z), slot PR_*, index n (wgsl2d :92-229). */
static inline REAL SFX(pr)(const OracleParams *p, int axis, int slot, int n)
{
    int na = axis == 0 ? p->nx : (axis == 1 ? p->ny : p->nz);
    int pad = (p->ord - 1) * 2;
    if (n < 0 || n >= na - pad) return K(0.0);
    return (REAL)p->prof[axis][slot][n];
}

/* Zener constants for mechanism l, recomputed exactly as the reference does
 * per cell (wgsl2d :881-894, wgsl3d :1405-1418). */
typedef struct { REAL phi_p, phi_s, h_p, h_s, m_p, m_s; } SFX(ZenerK);
static SFX(ZenerK) SFX(zener)(const OracleParams *p, int l)
{
    SFX(ZenerK) z;
    int L = p->n_sls;
    REAL dt = (REAL)p->dt;
    REAL t_ep = (REAL)p->tau[0 * L + l], t_sp = (REAL)p->tau[1 * L + l];
    REAL t_es = (REAL)p->tau[2 * L + l], t_ss = (REAL)p->tau[3 * L + l];
    REAL inv_p = K(1.0) / t_sp, inv_s = K(1.0) / t_ss;
    REAL al_p = t_ep / t_sp, al_s = t_es / t_ss;
    z.phi_p = dt * (K(1.0) - al_p) * inv_p / (REAL)p->sum_alpha[0];
    z.phi_s = dt * (K(1.0) - al_s) * inv_s / (REAL)p->sum_alpha[1];
    z.h_p = K(0.5) * dt * inv_p;
    z.h_s = K(0.5) * dt * inv_s;
    z.m_p = K(1.0) / (K(1.0) + z.h_p);
    z.m_s = K(1.0) / (K(1.0) + z.h_s);
    return z;
}

/* One staggered derivative term: WGSL divides each term by d (wgsl2d :850),
 * Taichi multiplies by f32(1/d) (sim_taichi_elastic.py:154-157). */
static inline REAL SFX(term)(const OracleParams *p, int c, REAL a, REAL b, REAL d, REAL one_d)
{
    REAL C = (REAL)p->coefs[c];
    if (p->variant == ORC_VARIANT_TAICHI) return C * (a - b) * one_d;
    return C * (a - b) / d;
}

/* CPML psi recursion + stretched derivative (wgsl2d :854-861). */
static inline REAL SFX(cpml)(const OracleParams *p, REAL *psi, long idx, int axis,
                             int half, int n, REAL D)
{
    int sa = half ? PR_AH : PR_A, sb = half ? PR_BH : PR_B, sk = half ? PR_KH : PR_K;
    REAL pnew = SFX(pr)(p, axis, sb, n) * psi[idx] + SFX(pr)(p, axis, sa, n) * D;
    psi[idx] = pnew;
    return D / SFX(pr)(p, axis, sk, n) + pnew;
}

/* ======================================================================== */
/* 2D                                                                        */
/* ======================================================================== */
static void SFX(stress2d)(const OracleParams *p, SFX(OracleState) *st)
{
    const int nx = p->nx, ny = p->ny, last = p->ord - 1, off = p->ord - 1;
    const REAL dx = (REAL)p->dx, dy = (REAL)p->dy, dt = (REAL)p->dt;
    const REAL one_dx = (REAL)(float)(1.0f / p->dx), one_dy = (REAL)(float)(1.0f / p->dy);
    const int L = p->n_sls;
    SFX(ZenerK) zk[ORC_MAX_SLS];
    for (int l = 0; l < L; l++) zk[l] = SFX(zener)(p, l);
    const int taichi = p->variant == ORC_VARIANT_TAICHI;

#pragma omp parallel for num_threads(p->nthreads) schedule(static)
    for (int x = 0; x < nx; x++) {
        for (int y = 0; y < ny; y++) {
            long id = (long)y + (long)x * ny;
            /* normal stresses: x in D+ range, y in D- range (wgsl2d :842-846) */
            if (x >= last && x < nx - last - 1 && y >= last + 1 && y < ny - last) {
                REAL dvx = K(0.0), dvy = K(0.0);
                for (int c = 0; c < p->ord; c++) {
                    dvx += SFX(term)(p, c, SFX(g2)(p, st->vx, x + c + 1, y), SFX(g2)(p, st->vx, x - c, y), dx, one_dx);
                    dvy += SFX(term)(p, c, SFX(g2)(p, st->vy, x, y + c), SFX(g2)(p, st->vy, x, y - c - 1), dy, one_dy);
                }
                dvx = SFX(cpml)(p, st->psi[PSI_DVX_DX], id, 0, 1, x - off, dvx);
                dvy = SFX(cpml)(p, st->psi[PSI_DVY_DY], id, 1, 0, y - off, dvy);

                REAL rho_h = K(0.5) * (SFX(m2)(p, p->rho, x + 1, y) + SFX(m2)(p, p->rho, x, y));
                REAL cp_h = K(0.5) * (SFX(m2)(p, p->cp, x + 1, y) + SFX(m2)(p, p->cp, x, y));
                REAL csa = SFX(m2)(p, p->cs, x + 1, y), csb = SFX(m2)(p, p->cs, x, y);
                REAL cs_l = K(0.5) * (csa + csb);
                REAL cs_m = ((csa < csb ? csa : csb) == K(0.0)) ? K(0.0) : cs_l;
                REAL lambda = rho_h * (cp_h * cp_h - K(2.0) * cs_l * cs_l);
                REAL mu = rho_h * (cs_m * cs_m);
                REAL lp2m = lambda + K(2.0) * mu;
                REAL lpm = lambda + mu;

                if (taichi) {
                    /* sim_taichi_elastic.py:168-182 (elastic only) */
                    REAL div = dvx + dvy;
                    REAL dl = lambda * div;
                    st->sxx[id] += (dl + K(2.0) * mu * dvx) * dt;
                    st->syy[id] += (dl + K(2.0) * mu * dvy) * dt;
                } else if (p->visco) {
                    /* wgsl2d :874-912 */
                    REAL sxx_r = K(0.0), syy_r = K(0.0);
                    for (int l = 0; l < L; l++) {
                        REAL *rxx = &st->r[0][id * L + l], *ryy = &st->r[1][id * L + l];
                        REAL ro_xx = *rxx, ro_yy = *ryy;
                        REAL rn_xx = ((ro_xx + (dvx + dvy) * zk[l].phi_p - ro_xx * zk[l].h_p) * zk[l].m_p);
                        REAL rn_yy = ((ro_yy + K(0.5) * (dvx - dvy) * zk[l].phi_s - ro_yy * zk[l].h_s) * zk[l].m_s);
                        sxx_r += ro_xx + rn_xx;
                        syy_r += ro_yy + rn_yy;
                        *rxx = rn_xx;
                        *ryy = rn_yy;
                    }
                    st->sxx[id] = st->sxx[id] + (lp2m * dvx + lambda * dvy + K(0.5) * lpm * sxx_r + mu * syy_r) * dt;
                    st->syy[id] = st->syy[id] + (lambda * dvx + lp2m * dvy + K(0.5) * lpm * sxx_r - mu * syy_r) * dt;
                } else {
                    /* wgsl2d :915-916 */
                    st->sxx[id] = st->sxx[id] + (lp2m * dvx + lambda * dvy) * dt;
                    st->syy[id] = st->syy[id] + (lambda * dvx + lp2m * dvy) * dt;
                }
            }
            /* shear stress: x in D- range, y in D+ range (wgsl2d :927-931) */
            if (x >= last + 1 && x < nx - last && y >= last && y < ny - last - 1) {
                REAL dvydx = K(0.0), dvxdy = K(0.0);
                for (int c = 0; c < p->ord; c++) {
                    dvydx += SFX(term)(p, c, SFX(g2)(p, st->vy, x + c, y), SFX(g2)(p, st->vy, x - c - 1, y), dx, one_dx);
                    dvxdy += SFX(term)(p, c, SFX(g2)(p, st->vx, x, y + c + 1), SFX(g2)(p, st->vx, x, y - c), dy, one_dy);
                }
                dvydx = SFX(cpml)(p, st->psi[PSI_DVY_DX], id, 0, 0, x - off, dvydx);
                dvxdy = SFX(cpml)(p, st->psi[PSI_DVX_DY], id, 1, 1, y - off, dvxdy);

                REAL rho_h = K(0.5) * (SFX(m2)(p, p->rho, x, y + 1) + SFX(m2)(p, p->rho, x, y));
                REAL csa = SFX(m2)(p, p->cs, x, y + 1), csb = SFX(m2)(p, p->cs, x, y);
                REAL cs_h = ((csa < csb ? csa : csb) == K(0.0)) ? K(0.0) : K(0.5) * (csa + csb);
                REAL mu = rho_h * (cs_h * cs_h);
                if (p->visco && !taichi) {
                    /* wgsl2d :953-974 */
                    REAL sxy_r = K(0.0);
                    for (int l = 0; l < L; l++) {
                        REAL *rxy = &st->r[3][id * L + l];
                        REAL ro = *rxy;
                        REAL rn = ((ro + (dvydx + dvxdy) * zk[l].phi_s - ro * zk[l].h_s) * zk[l].m_s);
                        sxy_r += ro + rn;
                        *rxy = rn;
                    }
                    st->sxy[id] = st->sxy[id] + (mu * (dvydx + dvxdy) + K(0.5) * mu * sxy_r) * dt;
                } else {
                    /* wgsl2d :977; sim_taichi_elastic.py:229 */
                    st->sxy[id] = st->sxy[id] + (dvxdy + dvydx) * mu * dt;
                }
            }
        }
    }
}

static void SFX(velocity2d)(const OracleParams *p, SFX(OracleState) *st, int it)
{
    const int nx = p->nx, ny = p->ny, last = p->ord - 1, off = p->ord - 1;
    const REAL dx = (REAL)p->dx, dy = (REAL)p->dy, dt = (REAL)p->dt;
    const REAL one_dx = (REAL)(float)(1.0f / p->dx), one_dy = (REAL)(float)(1.0f / p->dy);
    const int taichi = p->variant == ORC_VARIANT_TAICHI;
    REAL vmax = *st->v2max;

#pragma omp parallel for num_threads(p->nthreads) schedule(static) reduction(max:vmax)
    for (int x = 0; x < nx; x++) {
        for (int y = 0; y < ny; y++) {
            long id = (long)y + (long)x * ny;
            int in_vx = x >= last + 1 && x < nx - last && y >= last + 1 && y < ny - last;
            int in_vy = x >= last && x < nx - last - 1 && y >= last && y < ny - last - 1;
            /* vx (wgsl2d :998-1025; sim_taichi_elastic.py:240-268) */
            if (in_vx) {
                REAL dsxx = K(0.0), dsxy = K(0.0);
                for (int c = 0; c < p->ord; c++) {
                    dsxx += SFX(term)(p, c, SFX(g2)(p, st->sxx, x + c, y), SFX(g2)(p, st->sxx, x - c - 1, y), dx, one_dx);
                    dsxy += SFX(term)(p, c, SFX(g2)(p, st->sxy, x, y + c), SFX(g2)(p, st->sxy, x, y - c - 1), dy, one_dy);
                }
                dsxx = SFX(cpml)(p, st->psi[PSI_DSXX_DX], id, 0, 0, x - off, dsxx);
                dsxy = SFX(cpml)(p, st->psi[PSI_DSXY_DY], id, 1, 0, y - off, dsxy);
                REAL rho = SFX(m2)(p, p->rho, x, y);
                if (rho > K(0.0)) {
                    if (taichi) st->vx[id] += (dsxx + dsxy) * dt / rho;
                    else st->vx[id] = (dsxx + dsxy) * dt / rho + st->vx[id];
                }
            } else if (taichi) {
                st->vx[id] = K(0.0);
            }
            /* vy (wgsl2d :1027-1065; sim_taichi_elastic.py:270-304) */
            REAL rho;
            if (taichi)
                rho = K(0.25) * (SFX(m2)(p, p->rho, x, y + 1) + SFX(m2)(p, p->rho, x, y)
                                 + SFX(m2)(p, p->rho, x + 1, y + 1) + SFX(m2)(p, p->rho, x + 1, y));
            else
                rho = K(0.25) * (SFX(m2)(p, p->rho, x, y) + SFX(m2)(p, p->rho, x + 1, y)
                                 + SFX(m2)(p, p->rho, x + 1, y + 1) + SFX(m2)(p, p->rho, x, y + 1));
            if (in_vy) {
                REAL dsxy = K(0.0), dsyy = K(0.0);
                for (int c = 0; c < p->ord; c++) {
                    dsxy += SFX(term)(p, c, SFX(g2)(p, st->sxy, x + c + 1, y), SFX(g2)(p, st->sxy, x - c, y), dx, one_dx);
                    dsyy += SFX(term)(p, c, SFX(g2)(p, st->syy, x, y + c + 1), SFX(g2)(p, st->syy, x, y - c), dy, one_dy);
                }
                dsxy = SFX(cpml)(p, st->psi[PSI_DSXY_DX], id, 0, 1, x - off, dsxy);
                dsyy = SFX(cpml)(p, st->psi[PSI_DSYY_DY], id, 1, 1, y - off, dsyy);
                if (rho > K(0.0)) {
                    if (taichi) {
                        int is = p->src_map[id];
                        REAL vs = K(0.0);
                        if (is != -1) vs = (REAL)p->src[(long)(it - 1) * p->src_stride + is] * dt / rho;
                        st->vy[id] += vs + (dsxy + dsyy) * dt / rho;
                    } else {
                        st->vy[id] = (dsxy + dsyy) * dt / rho + st->vy[id];
                    }
                }
            } else {
                if (!taichi) st->vx[id] = K(0.0);
                st->vy[id] = K(0.0);
            }
            if (!taichi) {
                /* source after the update, anywhere (wgsl2d :1059-1065); device
                 * counter it is 0-based and get_source_term is bounds-checked
                 * against (n_iter, n_src_el) (wgsl2d :67-71) */
                int is = p->src_map[id];
                if (is != -1 && rho > K(0.0)) {
                    REAL val = K(0.0);
                    if (it >= 0 && it < p->n_iter && is >= 0 && is < p->src_stride)
                        val = (REAL)p->src[(long)is + (long)it * p->src_stride];
                    st->vy[id] = st->vy[id] + val * dt / rho;
                }
            }
            REAL v2 = st->vx[id] * st->vx[id] + st->vy[id] * st->vy[id];
            if (!(v2 <= vmax)) vmax = v2;
        }
    }
    *st->v2max = vmax;
}

/* Receivers: WGSL table loop (wgsl2d :1117-1134, wgsl3d :1823-1843). */
/* The field receivers sum: the reference's sigma_yy (2D) / sigma_zz (3D)
 * (wgsl2d :1130, wgsl3d :1839), or the channel a test asks for. */
static const REAL *SFX(rec_field)(const OracleParams *p, const SFX(OracleState) *st)
{
    REAL *const f[9] = {st->vx, st->vy, st->vz, st->sxx, st->syy, st->szz, st->sxy, st->sxz, st->syz};
    return f[p->rec_channel];
}

static int SFX(velocity_channel)(const OracleParams *p)
{
    return p->rec_channel == ORC_CH_VX || p->rec_channel == ORC_CH_VY || p->rec_channel == ORC_CH_VZ;
}

static void SFX(sensors_wgsl)(const OracleParams *p, SFX(OracleState) *st, int it)
{
    const int cols = p->ndim + 1;
    const REAL *f = SFX(rec_field)(p, st);
    for (int s = 0; s < p->n_rec_el; s++) {
        int pt = p->offset_sensors[s];
        for (;;) {
            if (pt < 0 || pt >= p->n_rec_pt) break;
            if (p->info_rec_pt[(long)pt * cols + (cols - 1)] != s) break;
            if (it >= p->delay_rec[s]) {
                int x = p->info_rec_pt[(long)pt * cols + 0];
                int y = p->info_rec_pt[(long)pt * cols + 1];
                REAL val = p->ndim == 2 ? SFX(g2)(p, f, x, y)
                                        : SFX(g3)(p, f, x, y, p->info_rec_pt[(long)pt * cols + 2]);
                if (it >= 0 && it < p->n_iter) {
                    REAL *t = &st->traces[(long)it * p->n_rec_el + s];
                    *t = *t + val;
                }
            }
            pt++;
        }
    }
}

/* Taichi receivers: inside the normal-stress block, row it-1, host it
 * 1-based (sim_taichi_elastic.py:187-190).  Cells are visited in C order; the
 * reference's atomic += has no defined order. */
static void SFX(sensors_taichi)(const OracleParams *p, SFX(OracleState) *st, int it)
{
    const int nx = p->nx, ny = p->ny, last = p->ord - 1;
    const REAL *f = SFX(rec_field)(p, st);
    if (it < 1 || it > p->n_iter) return;
    for (int x = last; x < nx - last - 1; x++)
        for (int y = last + 1; y < ny - last; y++) {
            long id = (long)y + (long)x * ny;
            int s = p->rec_map[id];
            if (s >= 0 && s < p->n_rec_el && it >= p->delay_rec[s])
                st->traces[(long)(it - 1) * p->n_rec_el + s] += f[id];
        }
}

/* ======================================================================== */
/* 3D (WGSL)                                                                 */
/* ======================================================================== */
#define G3(f, x, y, z) SFX(g3)(p, (f), (x), (y), (z))
#define M3(f, x, y, z) SFX(m3)(p, (f), (x), (y), (z))

static void SFX(stress3d)(const OracleParams *p, SFX(OracleState) *st)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz, last = p->ord - 1, off = p->ord - 1;
    const REAL dx = (REAL)p->dx, dy = (REAL)p->dy, dz = (REAL)p->dz, dt = (REAL)p->dt;
    const int L = p->n_sls;
    SFX(ZenerK) zk[ORC_MAX_SLS];
    for (int l = 0; l < L; l++) zk[l] = SFX(zener)(p, l);

#pragma omp parallel for num_threads(p->nthreads) schedule(static)
    for (int x = 0; x < nx; x++)
    for (int y = 0; y < ny; y++)
    for (int z = 0; z < nz; z++) {
        long id = SFX(i3)(p, x, y, z);
        /* normal stresses (wgsl3d :1356-1453) */
        if (x >= last && x < nx - last - 1 && y >= last + 1 && y < ny - last && z >= last + 1 && z < nz - last) {
            REAL dvx = K(0.0), dvy = K(0.0), dvz = K(0.0);
            for (int c = 0; c < p->ord; c++) {
                REAL C = (REAL)p->coefs[c];
                dvx += C * (G3(st->vx, x + c + 1, y, z) - G3(st->vx, x - c, y, z)) / dx;
                dvy += C * (G3(st->vy, x, y + c, z) - G3(st->vy, x, y - c - 1, z)) / dy;
                dvz += C * (G3(st->vz, x, y, z + c) - G3(st->vz, x, y, z - c - 1)) / dz;
            }
            dvx = SFX(cpml)(p, st->psi[PSI_DVX_DX], id, 0, 1, x - off, dvx);
            dvy = SFX(cpml)(p, st->psi[PSI_DVY_DY], id, 1, 0, y - off, dvy);
            dvz = SFX(cpml)(p, st->psi[PSI_DVZ_DZ], id, 2, 0, z - off, dvz);
            REAL rho = M3(p->rho, x, y, z), cp = M3(p->cp, x, y, z), cs = M3(p->cs, x, y, z);
            REAL lambda = rho * (cp * cp - K(2.0) * cs * cs);
            REAL mu = rho * (cs * cs);
            REAL lp2m = lambda + K(2.0) * mu;
            if (p->visco) {
                REAL Sxx = K(0.0), Syy = K(0.0), Szz = K(0.0);
                REAL div = dvx + dvy + dvz;
                for (int l = 0; l < L; l++) {
                    REAL *pxx = &st->r[0][id * L + l], *pyy = &st->r[1][id * L + l], *pzz = &st->r[2][id * L + l];
                    REAL oxx = *pxx, oyy = *pyy, ozz = *pzz;
                    REAL nxx = ((oxx + div * zk[l].phi_p - oxx * zk[l].h_p) * zk[l].m_p);
                    REAL nyy = ((oyy + (dvx - div / K(3.0)) * zk[l].phi_s - oyy * zk[l].h_s) * zk[l].m_s);
                    REAL nzz = ((ozz + (dvy - div / K(3.0)) * zk[l].phi_s - ozz * zk[l].h_s) * zk[l].m_s);
                    Sxx += oxx + nxx; Syy += oyy + nyy; Szz += ozz + nzz;
                    *pxx = nxx; *pyy = nyy; *pzz = nzz;
                }
                REAL l23 = lambda + K(2.0) * mu / K(3.0);
                st->sxx[id] = st->sxx[id] + (lp2m * dvx + lambda * (dvy + dvz) + l23 * K(0.5) * Sxx + mu * Syy) * dt;
                st->syy[id] = st->syy[id] + (lambda * (dvx + dvz) + lp2m * dvy + l23 * K(0.5) * Sxx + mu * Szz) * dt;
                st->szz[id] = st->szz[id] + (lambda * (dvx + dvy) + lp2m * dvz + lp2m * K(0.5) * Sxx - (mu / K(3.0)) * (Syy + Szz)) * dt;
            } else {
                st->sxx[id] = st->sxx[id] + (lp2m * dvx + lambda * (dvy + dvz)) * dt;
                st->syy[id] = st->syy[id] + (lambda * (dvx + dvz) + lp2m * dvy) * dt;
                st->szz[id] = st->szz[id] + (lambda * (dvx + dvy) + lp2m * dvz) * dt;
            }
        }
        /* shear stresses (wgsl3d :1457-1629) */
        for (int sh = 0; sh < 3; sh++) {
            int ok;
            if (sh == 0)      ok = x >= last + 1 && x < nx - last && y >= last && y < ny - last - 1 && z >= last && z < nz - last - 1;
            else if (sh == 1) ok = x >= last + 1 && x < nx - last && y >= last && y < ny - last - 1 && z >= last && z < nz - last - 1;
            else              ok = x >= last && x < nx - last - 1 && y >= last && y < ny - last - 1 && z >= last && z < nz - last - 1;
            if (!ok) continue;
            REAL d1 = K(0.0), d2 = K(0.0), rho, cs;
            if (sh == 0) {          /* sigma_xy: dvy/dx (D-), dvx/dy (D+) */
                for (int c = 0; c < p->ord; c++) {
                    REAL C = (REAL)p->coefs[c];
                    d1 += C * (G3(st->vy, x + c, y, z) - G3(st->vy, x - c - 1, y, z)) / dx;
                    d2 += C * (G3(st->vx, x, y + c + 1, z) - G3(st->vx, x, y - c, z)) / dy;
                }
                d1 = SFX(cpml)(p, st->psi[PSI_DVY_DX], id, 0, 0, x - off, d1);
                d2 = SFX(cpml)(p, st->psi[PSI_DVX_DY], id, 1, 1, y - off, d2);
                rho = K(0.25) * (M3(p->rho, x + 1, y, z) + M3(p->rho, x, y, z) + M3(p->rho, x, y + 1, z) + M3(p->rho, x + 1, y + 1, z));
                cs = K(0.25) * (M3(p->cs, x + 1, y, z) + M3(p->cs, x, y, z) + M3(p->cs, x, y + 1, z) + M3(p->cs, x + 1, y + 1, z));
            } else if (sh == 1) {   /* sigma_xz: dvz/dx (D-), dvx/dz (D+) */
                for (int c = 0; c < p->ord; c++) {
                    REAL C = (REAL)p->coefs[c];
                    d1 += C * (G3(st->vz, x + c, y, z) - G3(st->vz, x - c - 1, y, z)) / dx;
                    d2 += C * (G3(st->vx, x, y, z + c + 1) - G3(st->vx, x, y, z - c)) / dz;
                }
                d1 = SFX(cpml)(p, st->psi[PSI_DVZ_DX], id, 0, 0, x - off, d1);
                d2 = SFX(cpml)(p, st->psi[PSI_DVX_DZ], id, 2, 1, z - off, d2);
                rho = K(0.25) * (M3(p->rho, x + 1, y, z) + M3(p->rho, x, y, z) + M3(p->rho, x, y, z + 1) + M3(p->rho, x + 1, y, z + 1));
                cs = K(0.25) * (M3(p->cs, x + 1, y, z) + M3(p->cs, x, y, z) + M3(p->cs, x, y, z + 1) + M3(p->cs, x + 1, y, z + 1));
            } else {                /* sigma_yz: dvz/dy (D+), dvy/dz (D+) */
                for (int c = 0; c < p->ord; c++) {
                    REAL C = (REAL)p->coefs[c];
                    d1 += C * (G3(st->vz, x, y + c + 1, z) - G3(st->vz, x, y - c, z)) / dy;
                    d2 += C * (G3(st->vy, x, y, z + c + 1) - G3(st->vy, x, y, z - c)) / dz;
                }
                d1 = SFX(cpml)(p, st->psi[PSI_DVZ_DY], id, 1, 1, y - off, d1);
                d2 = SFX(cpml)(p, st->psi[PSI_DVY_DZ], id, 2, 1, z - off, d2);
                rho = K(0.25) * (M3(p->rho, x, y + 1, z) + M3(p->rho, x, y, z) + M3(p->rho, x, y, z + 1) + M3(p->rho, x, y + 1, z + 1));
                cs = K(0.25) * (M3(p->cs, x, y + 1, z) + M3(p->cs, x, y, z) + M3(p->cs, x, y, z + 1) + M3(p->cs, x, y + 1, z + 1));
            }
            REAL mu = rho * (cs * cs);
            REAL *sig = sh == 0 ? st->sxy : (sh == 1 ? st->sxz : st->syz);
            if (p->visco) {
                REAL S = K(0.0);
                REAL X = d1 + d2;   /* (dvy_dx+dvx_dy), (dvz_dx+dvx_dz), (dvz_dy+dvy_dz) */
                for (int l = 0; l < L; l++) {
                    REAL *pr_ = &st->r[3 + sh][id * L + l];
                    REAL o = *pr_;
                    REAL n_ = (o + X * zk[l].phi_s - o * zk[l].h_s) * zk[l].m_s;
                    S += o + n_;
                    *pr_ = n_;
                }
                /* sigma_xy: mu*(dvy_dx+dvx_dy); sigma_xz: mu*(dvx_dz+dvz_dx);
                 * sigma_yz: mu*(dvy_dz+dvz_dy) -- additions commute exactly */
                sig[id] = sig[id] + (mu * (d1 + d2) + mu * K(0.5) * S) * dt;
            } else {
                sig[id] = sig[id] + (d2 + d1) * mu * dt;
            }
        }
    }
}

static void SFX(velocity3d)(const OracleParams *p, SFX(OracleState) *st, int it)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz, last = p->ord - 1, off = p->ord - 1;
    const REAL dx = (REAL)p->dx, dy = (REAL)p->dy, dz = (REAL)p->dz, dt = (REAL)p->dt;
    REAL vmax = *st->v2max;

#pragma omp parallel for num_threads(p->nthreads) schedule(static) reduction(max:vmax)
    for (int x = 0; x < nx; x++)
    for (int y = 0; y < ny; y++)
    for (int z = 0; z < nz; z++) {
        long id = SFX(i3)(p, x, y, z);
        /* vx (wgsl3d :1647-1681) */
        if (x >= last + 1 && x < nx - last && y >= last + 1 && y < ny - last && z >= last + 1 && z < nz - last) {
            REAL a = K(0.0), b = K(0.0), c_ = K(0.0);
            for (int c = 0; c < p->ord; c++) {
                REAL C = (REAL)p->coefs[c];
                a += C * (G3(st->sxx, x + c, y, z) - G3(st->sxx, x - c - 1, y, z)) / dx;
                b += C * (G3(st->sxy, x, y + c, z) - G3(st->sxy, x, y - c - 1, z)) / dy;
                c_ += C * (G3(st->sxz, x, y, z + c) - G3(st->sxz, x, y, z - c - 1)) / dz;
            }
            a = SFX(cpml)(p, st->psi[PSI_DSXX_DX], id, 0, 0, x - off, a);
            b = SFX(cpml)(p, st->psi[PSI_DSXY_DY], id, 1, 0, y - off, b);
            c_ = SFX(cpml)(p, st->psi[PSI_DSXZ_DZ], id, 2, 0, z - off, c_);
            REAL rho = K(0.5) * (M3(p->rho, x + 1, y, z) + M3(p->rho, x, y, z));
            if (rho > K(0.0)) st->vx[id] = (a + b + c_) * dt / rho + st->vx[id];
        }
        /* vy (wgsl3d :1683-1717); its density also serves vz and the source */
        REAL rho = K(0.5) * (M3(p->rho, x, y + 1, z) + M3(p->rho, x, y, z));
        if (x >= last && x < nx - last - 1 && y >= last && y < ny - last - 1 && z >= last + 1 && z < nz - last) {
            REAL a = K(0.0), b = K(0.0), c_ = K(0.0);
            for (int c = 0; c < p->ord; c++) {
                REAL C = (REAL)p->coefs[c];
                a += C * (G3(st->sxy, x + c + 1, y, z) - G3(st->sxy, x - c, y, z)) / dx;
                b += C * (G3(st->syy, x, y + c + 1, z) - G3(st->syy, x, y - c, z)) / dy;
                c_ += C * (G3(st->syz, x, y, z + c) - G3(st->syz, x, y, z - c - 1)) / dz;
            }
            a = SFX(cpml)(p, st->psi[PSI_DSXY_DX], id, 0, 1, x - off, a);
            b = SFX(cpml)(p, st->psi[PSI_DSYY_DY], id, 1, 1, y - off, b);
            c_ = SFX(cpml)(p, st->psi[PSI_DSYZ_DZ], id, 2, 0, z - off, c_);
            if (rho > K(0.0)) st->vy[id] = (a + b + c_) * dt / rho + st->vy[id];
        }
        /* vz (wgsl3d :1719-1758) */
        if (x >= last && x < nx - last - 1 && y >= last + 1 && y < ny - last && z >= last && z < nz - last - 1) {
            REAL a = K(0.0), b = K(0.0), c_ = K(0.0);
            for (int c = 0; c < p->ord; c++) {
                REAL C = (REAL)p->coefs[c];
                a += C * (G3(st->sxz, x + c + 1, y, z) - G3(st->sxz, x - c, y, z)) / dx;
                b += C * (G3(st->syz, x, y + c, z) - G3(st->syz, x, y - c - 1, z)) / dy;
                c_ += C * (G3(st->szz, x, y, z + c + 1) - G3(st->szz, x, y, z - c)) / dz;
            }
            a = SFX(cpml)(p, st->psi[PSI_DSXZ_DX], id, 0, 1, x - off, a);
            b = SFX(cpml)(p, st->psi[PSI_DSYZ_DY], id, 1, 0, y - off, b);
            c_ = SFX(cpml)(p, st->psi[PSI_DSZZ_DZ], id, 2, 1, z - off, c_);
            if (rho > K(0.0)) st->vz[id] = (a + b + c_) * dt / rho + st->vz[id];
        } else {
            st->vx[id] = K(0.0); st->vy[id] = K(0.0); st->vz[id] = K(0.0);
        }
        /* source into vz with the vy density (wgsl3d :1760-1766) */
        int is = p->src_map[id];
        if (is != -1 && rho > K(0.0)) {
            REAL val = K(0.0);
            if (it >= 0 && it < p->n_iter && is >= 0 && is < p->src_stride)
                val = (REAL)p->src[(long)is + (long)it * p->src_stride];
            st->vz[id] = st->vz[id] + val * dt / rho;
        }
        REAL v2 = st->vx[id] * st->vx[id] + st->vy[id] * st->vy[id] + st->vz[id] * st->vz[id];
        if (!(v2 <= vmax)) vmax = v2;
    }
    *st->v2max = vmax;
}
#undef G3
#undef M3

/* ---- driver: run steps [s0, s0+n) of the reference loop -------------------
 * s is the 0-based step index: host it = s+1 (sim_webgpu_viscoelastic.py:471),
 * device counter it = s (param slot 7 starts at 0 and is incremented after
 * each step, wgsl2d :1136-1141). */
int SFX(oracle_run)(const OracleParams *p, SFX(OracleState) *st, int s0, int n)
{
    for (int s = s0; s < s0 + n; s++) {
        if (p->ndim == 2) {
            SFX(stress2d)(p, st);
            if (p->variant == ORC_VARIANT_TAICHI) {
                /* the reference gathers sigma_yy inside the stress pass; a
                 * velocity channel is read after the velocity pass */
                if (!SFX(velocity_channel)(p)) SFX(sensors_taichi)(p, st, s + 1);
                SFX(velocity2d)(p, st, s + 1);
                if (SFX(velocity_channel)(p)) SFX(sensors_taichi)(p, st, s + 1);
            } else {
                SFX(velocity2d)(p, st, s);
                SFX(sensors_wgsl)(p, st, s);
            }
        } else {
            SFX(stress3d)(p, st);
            SFX(velocity3d)(p, st, s);
            SFX(sensors_wgsl)(p, st, s);
        }
    }
    return 0;
}

#undef K
