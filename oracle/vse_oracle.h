/*
 * vse_oracle.h -- CPU restatement of the reference's staggered-grid
 * velocity-stress time loop.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker (or as the
 * timed CPU baseline).  The product path (taichi_sim_amd) never calls it.
 *
 * Every update follows the reference text literally, including its bounds-
 * checked accessors (an out-of-range read returns 0):
 *   2D WGSL   : shader_webgpu_viscoelastic.wgsl:829-1072 (stress, velocity),
 *               :1117-1141 (receivers, step counter)
 *   2D Taichi : sim_taichi_elastic.py:138-309
 *   3D WGSL   : shader_webgpu_3D_viscoelastic.wgsl:1340-1773, :1823-1850
 * The same body is compiled twice: REAL = float (FP32 oracle) and
 * REAL = double (FP64 oracle).  Inputs are always FP32 arrays, promoted.
 */
#ifndef VSE_ORACLE_H
#define VSE_ORACLE_H

#define ORC_VARIANT_WGSL   0
#define ORC_VARIANT_TAICHI 1
#define ORC_MAX_ORD 8
#define ORC_MAX_SLS 8

/* Memory-variable (CPML psi) slots, in the reference's naming. */
enum {
    /* stress pass: derivatives of v */
    PSI_DVX_DX = 0, PSI_DVY_DY, PSI_DVZ_DZ, PSI_DVY_DX, PSI_DVX_DY,
    PSI_DVZ_DX, PSI_DVX_DZ, PSI_DVZ_DY, PSI_DVY_DZ,
    /* velocity pass: derivatives of sigma */
    PSI_DSXX_DX, PSI_DSXY_DY, PSI_DSXZ_DZ, PSI_DSXY_DX, PSI_DSYY_DY,
    PSI_DSYZ_DZ, PSI_DSXZ_DX, PSI_DSYZ_DY, PSI_DSZZ_DZ,
    PSI_COUNT
};

/* CPML profile slots per axis: full grid a,b,k then half grid a,b,k. */
enum { PR_A = 0, PR_B, PR_K, PR_AH, PR_BH, PR_KH };

typedef struct {
    int ndim;               /* 2 or 3 */
    int nx, ny, nz;         /* full grid (2D: nz ignored, y is the ROI z axis) */
    int ord;                /* number of FD coefficients */
    float coefs[ORC_MAX_ORD];
    float dx, dy, dz, dt;
    int variant;            /* ORC_VARIANT_* (3D: WGSL only) */
    int visco;              /* viscoelastic_attn flag */
    int n_sls;              /* Zener mechanisms L */
    float tau[4 * ORC_MAX_SLS];   /* rows eps_p, sig_p, eps_s, sig_s; (4, L) row-major */
    float sum_alpha[2];
    const float *rho, *cp, *cs;   /* full grid, C order */
    const float *prof[3][6];      /* per axis (x,y,z): a,b,k,a_h,b_h,k_h; length n-2*(ord-1) */
    /* sources */
    const float *src;       /* flat source buffer as uploaded by the reference */
    int n_iter;             /* rows of the source / trace buffers */
    int src_stride;         /* WGSL: n_src_el (row stride used by get_source_term);
                               Taichi: number of source columns */
    const int *src_map;     /* full grid, -1 or source column */
    /* receivers (WGSL table form; Taichi uses rec_map) */
    int n_rec_el;           /* receiver elements (trace columns) */
    int n_rec_pt;           /* receiver points */
    const int *info_rec_pt; /* (n_rec_pt, ndim+1): coords then id */
    const int *offset_sensors; /* (n_rec_el) */
    const int *delay_rec;   /* (n_rec_el) */
    const int *rec_map;     /* Taichi: full grid, -1 or receiver id */
    int rec_channel;        /* ORC_CH_*: the field the receivers sum (the reference's: sigma_yy / sigma_zz) */
    int nthreads;
} OracleParams;

/* Receiver channels (field order of the state). */
enum { ORC_CH_VX = 0, ORC_CH_VY, ORC_CH_VZ, ORC_CH_SXX, ORC_CH_SYY, ORC_CH_SZZ, ORC_CH_SXY, ORC_CH_SXZ, ORC_CH_SYZ };

#endif
