/*
 * CPU restatement of the compas-toolkit simulation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: it is compiled twice by
 * oracle.c, once with REAL = double (the FP64 oracle, suffix _f64) and once
 * with REAL = float (the FP32 port of the reference algorithm, suffix _f32,
 * also timed as the CPU baseline).  Only tests, the smoke check and the
 * benchmark's CPU-baseline leg may load it; the product library never does.
 *
 * Every function names the reference file:line range whose semantics it
 * restates.  Array conventions follow the reference: row-major, last index
 * fastest, complex values interleaved (re, im).  Tissue, sequence and
 * trajectory inputs are float32 because the reference stores them as float;
 * magnetisation inputs are REAL so that an FP64 chain can be evaluated.
 *
 * Requires: REAL, SFX(name) defined by the includer.
 */

#ifndef REAL
#error "define REAL before including oracle_impl.h"
#endif

typedef struct {
    REAL re, im;
} SFX(cpx);

#define CPX SFX(cpx)

static inline CPX SFX(mk)(REAL re, REAL im) {
    CPX r;
    r.re = re;
    r.im = im;
    return r;
}
static inline CPX SFX(cadd)(CPX a, CPX b) {
    return SFX(mk)(a.re + b.re, a.im + b.im);
}
static inline CPX SFX(csub)(CPX a, CPX b) {
    return SFX(mk)(a.re - b.re, a.im - b.im);
}
static inline CPX SFX(cmul)(CPX a, CPX b) {
    return SFX(mk)(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline CPX SFX(cscale)(CPX a, REAL s) {
    return SFX(mk)(a.re * s, a.im * s);
}
static inline CPX SFX(cconj)(CPX a) {
    return SFX(mk)(a.re, -a.im);
}
/* exp(z) = polar(exp(re), im)  -- complex_type.h:106-114 */
static inline CPX SFX(cexp)(CPX z) {
    REAL e = EXP(z.re);
    return SFX(mk)(e * COS(z.im), e * SIN(z.im));
}
/* polar(1, angle) -- complex_type.h:89-99 */
static inline CPX SFX(cpolar1)(REAL angle) {
    return SFX(mk)(COS(angle), SIN(angle));
}
static inline CPX SFX(ldc)(const float* p, long i) {
    return SFX(mk)((REAL)p[2 * i], (REAL)p[2 * i + 1]);
}
static inline CPX SFX(ldr)(const REAL* p, long i) {
    return SFX(mk)(p[2 * i], p[2 * i + 1]);
}
static inline void SFX(st)(REAL* p, long i, CPX v) {
    p[2 * i] = v.re;
    p[2 * i + 1] = v.im;
}

#define C_MK    SFX(mk)
#define C_ADD   SFX(cadd)
#define C_SUB   SFX(csub)
#define C_MUL   SFX(cmul)
#define C_SCALE SFX(cscale)
#define C_CONJ  SFX(cconj)
#define C_EXP   SFX(cexp)
#define C_POLAR SFX(cpolar1)
#define C_LDC   SFX(ldc)
#define C_LDR   SFX(ldr)
#define C_ST    SFX(st)

/* ------------------------------------------------------------------------ */
/* FISP extended phase graph                                                 */
/* ------------------------------------------------------------------------ */

/* EPGExciteMatrix(RF, B1): epg.cuh:36-90 */
typedef struct {
    REAL R11;
    CPX R12, R13;
    REAL R33;
} SFX(excite);

static inline SFX(excite) SFX(make_excite)(CPX RF, REAL B1) {
    SFX(excite) m;
    REAL abs_RF = HYPOT(RF.re, RF.im);
    REAL alpha = ((REAL)M_PI / (REAL)180) * abs_RF * B1;
    REAL x = alpha * (REAL)0.5;
    REAL sinx = SIN(x), cosx = COS(x);
    REAL sinx_sq = sinx * sinx, cosx_sq = cosx * cosx;
    REAL sinalpha = (REAL)2 * sinx * cosx;
    REAL cosalpha = (REAL)2 * cosx_sq - (REAL)1;
    REAL cosphi = 1, sinphi = 0;
    if (abs_RF > (REAL)FLT_MIN) {
        cosphi = RF.re / abs_RF;
        sinphi = RF.im / abs_RF;
    }
    REAL sin2phi = ((REAL)2 * cosphi) * sinphi;
    REAL cos2phi = ((REAL)2 * cosphi) * cosphi - (REAL)1;
    m.R11 = cosx_sq;
    m.R12 = C_MK(cos2phi * sinx_sq, sin2phi * sinx_sq);
    m.R13 = C_MK(sinphi * sinalpha, -cosphi * sinalpha);
    m.R33 = cosalpha;
    return m;
}

/* EPGExciteMatrix::apply: epg.cuh:92-111 */
static inline void SFX(apply_excite)(const SFX(excite) * m, CPX* Fp, CPX* Fm, CPX* Z) {
    CPX R12 = m->R12, R13 = m->R13;
    CPX R21 = C_CONJ(R12), R23 = C_CONJ(R13);
    CPX R31 = C_MK((REAL)-0.5 * R13.re, (REAL)0.5 * R13.im);
    CPX R32 = C_CONJ(R31);
    CPX fp = *Fp, fm = *Fm, z = *Z;
    *Fp = C_ADD(C_ADD(C_SCALE(fp, m->R11), C_MUL(R12, fm)), C_MUL(R13, z));
    *Fm = C_ADD(C_ADD(C_SCALE(fm, m->R11), C_MUL(R21, fp)), C_MUL(R23, z));
    *Z = C_ADD(C_ADD(C_MUL(R31, fp), C_MUL(R32, fm)), C_SCALE(z, m->R33));
}

/* off_resonance_rotation / calculate_E: fisp_kernels.cuh:110-121 */
static inline CPX SFX(off_res)(REAL dt, REAL B0) {
    if (B0 == (REAL)0) return C_MK(1, 0);
    return C_POLAR((REAL)(2.0 * M_PI) * dt * B0);
}
static inline REAL SFX(relax)(REAL dt, REAL T) {
    return EXP(-dt * ((REAL)1 / T));
}

/*
 * One voxel, one slice profile: fisp_kernels.cuh:99-223 with the EPG state
 * operations of epg.cuh:114-322 (state k = global index; the reference's
 * lane/item split is a layout choice and not part of the semantics).
 * Samples F+[0] at TE of the last repetition into echo[n] (+=).
 */
static void SFX(fisp_voxel)(
    REAL T1, REAL T2, REAL B1, REAL B0,
    int nTR, const float* RF, const float* profile,
    REAL TR, REAL TE, REAL TW, REAL TI,
    int N, int reps, int inversion, int wait_spoil,
    CPX* Fp, CPX* Fm, CPX* Z, CPX* echo) {
    REAL E1_TE = SFX(relax)(TE, T1), E2_TE = SFX(relax)(TE, T2);
    REAL E1_TI = SFX(relax)(TI, T1), E2_TI = SFX(relax)(TI, T2);
    REAL E1_TRmTE = SFX(relax)(TR - TE, T1), E2_TRmTE = SFX(relax)(TR - TE, T2);
    REAL E1_W = SFX(relax)(TW, T1), E2_W = SFX(relax)(TW, T2);
    CPX r_TE = SFX(off_res)(TE, B0);
    CPX r_TRmTE = SFX(off_res)(TR - TE, B0);
    CPX r_TW = SFX(off_res)(TW, B0);
    int k, n, rep;

    /* initialize(): epg.cuh:123-133 */
    for (k = 0; k < N; k++) {
        Fp[k] = C_MK(0, 0);
        Fm[k] = C_MK(0, 0);
        Z[k] = C_MK(0, 0);
    }
    Z[0] = C_MK(1, 0);

    for (rep = 0; rep < reps; rep++) {
        int sample = (rep == reps - 1); /* fisp_kernels.cuh:216-222 */
        if (inversion) {                /* :146-152 -- invert() ignores B1 */
            for (k = 0; k < N; k++) {
                Z[k] = C_SCALE(Z[k], (REAL)-1);
                Fp[k] = C_MK(0, 0);
                Fm[k] = C_MK(0, 0);
            }
            for (k = 0; k < N; k++) Z[k] = C_SCALE(Z[k], E1_TI);
            (void)E2_TI; /* F states are zero after spoiling; decay of zero is zero */
            Z[0].re += (REAL)1 - E1_TI;
        }
        for (n = 0; n < nTR; n++) {
            CPX rf = C_MUL(C_LDC(profile, n), C_LDC(RF, n)); /* :38-40 */
            SFX(excite) m = SFX(make_excite)(rf, B1);
            CPX a, b;
            for (k = 0; k < N; k++) SFX(apply_excite)(&m, &Fp[k], &Fm[k], &Z[k]);
            /* rotate_decay(E1_TE, E2_TE, r_TE): epg.cuh:182-190 */
            a = C_SCALE(r_TE, E2_TE);
            b = C_CONJ(a);
            for (k = 0; k < N; k++) {
                Fp[k] = C_MUL(Fp[k], a);
                Fm[k] = C_MUL(Fm[k], b);
                Z[k] = C_SCALE(Z[k], E1_TE);
            }
            Z[0].re += (REAL)1 - E1_TE; /* regrowth: epg.cuh:192-200 */
            if (sample) echo[n] = C_ADD(echo[n], Fp[0]);
            a = C_SCALE(r_TRmTE, E2_TRmTE);
            b = C_CONJ(a);
            for (k = 0; k < N; k++) {
                Fp[k] = C_MUL(Fp[k], a);
                Fm[k] = C_MUL(Fm[k], b);
                Z[k] = C_SCALE(Z[k], E1_TRmTE);
            }
            Z[0].re += (REAL)1 - E1_TRmTE;
            /* dephasing = shift_down + shift_up: epg.cuh:229-298 */
            for (k = 0; k + 1 < N; k++) Fm[k] = Fm[k + 1];
            Fm[N - 1] = C_MK(0, 0);
            for (k = N - 1; k >= 1; k--) Fp[k] = Fp[k - 1];
            Fp[0] = C_CONJ(Fm[0]);
        }
        if (wait_spoil) { /* :195-200 */
            for (k = 0; k < N; k++) {
                Fp[k] = C_MK(0, 0);
                Fm[k] = C_MK(0, 0);
            }
        }
        CPX aw = C_SCALE(r_TW, E2_W);
        CPX bw = C_CONJ(aw);
        for (k = 0; k < N; k++) {
            Fp[k] = C_MUL(Fp[k], aw);
            Fm[k] = C_MUL(Fm[k], bw);
            Z[k] = C_SCALE(Z[k], E1_W);
        }
        Z[0].re += (REAL)1 - E1_W;
    }
}

/*
 * simulate_magnetization(FISP): fisp.cu:7-43 (slice-profile loop summed into
 * zero-filled output), fisp.cu:105-140 and expand_undersampled_echos
 * (fisp_kernels.cuh:246-259): echos[r][v] = echo[v][r / usf].
 * echos: [nTR*usf][nv] complex.
 */
void SFX(fisp_simulate)(
    int nv, const float* T1, const float* T2, const float* B1, const float* B0,
    int nTR, const float* RF, int nslices, const float* profiles,
    float TR, float TE, float TW, float TI,
    int max_state, int usf, int reps, int inversion, int wait_spoil,
    REAL* echos, int nthreads) {
    int nr = nTR * usf;
    int N = max_state;
    long v;
#pragma omp parallel num_threads(nthreads)
    {
        CPX* Fp = (CPX*)malloc(sizeof(CPX) * (size_t)N);
        CPX* Fm = (CPX*)malloc(sizeof(CPX) * (size_t)N);
        CPX* Zs = (CPX*)malloc(sizeof(CPX) * (size_t)N);
        CPX* echo = (CPX*)malloc(sizeof(CPX) * (size_t)nTR);
#pragma omp for schedule(dynamic, 16)
        for (v = 0; v < nv; v++) {
            int j, n, u;
            for (n = 0; n < nTR; n++) echo[n] = C_MK(0, 0);
            for (j = 0; j < nslices; j++) {
                SFX(fisp_voxel)(
                    (REAL)T1[v], (REAL)T2[v], (REAL)B1[v], (REAL)B0[v],
                    nTR, RF, profiles + 2L * j * nTR,
                    (REAL)TR, (REAL)TE, (REAL)TW, (REAL)TI,
                    N, reps, inversion, wait_spoil, Fp, Fm, Zs, echo);
            }
            for (n = 0; n < nTR; n++)
                for (u = 0; u < usf; u++) C_ST(echos, (long)(n * usf + u) * nv + v, echo[n]);
        }
        free(Fp);
        free(Fm);
        free(Zs);
        free(echo);
    }
    (void)nr;
}

/* ------------------------------------------------------------------------ */
/* pSSFP isochromat simulation                                               */
/* ------------------------------------------------------------------------ */

typedef struct {
    REAL x, y, z;
} SFX(v3);
#define V3 SFX(v3)

/* Isochromat::rotate(gamma_dt_RF, gamma_dt_GR, r, dt, p): isochromat.cuh:14-73 */
static inline V3 SFX(iso_rotate_rf)(V3 m, CPX zap, REAL g, REAL r, REAL dt, REAL B1, REAL B0) {
    REAL ax = zap.re, ay = -zap.im, az = -(g * r);
    ax *= B1;
    ay *= B1;
    az -= dt * (REAL)(2.0 * M_PI) * B0;
    REAL theta = SQRT(ax * ax + ay * ay + az * az);
    if (FABS(theta) > (REAL)1e-9) {
        REAL kx = ax / theta, ky = ay / theta, kz = az / theta;
        REAL s = SIN(theta), c = COS(theta);
        REAL cx = ky * m.z - kz * m.y;
        REAL cy = kz * m.x - kx * m.z;
        REAL cz = kx * m.y - ky * m.x;
        REAL d = kx * m.x + ky * m.y + kz * m.z;
        REAL t = d * ((REAL)1 - c);
        V3 o;
        o.x = c * m.x + s * cx + t * kx;
        o.y = c * m.y + s * cy + t * ky;
        o.z = c * m.z + s * cz + t * kz;
        return o;
    }
    return m;
}

/* Isochromat::rotate(gamma_dt_GR, r, dt, p): isochromat.cuh:75-97 */
static inline V3 SFX(iso_rotate_z)(V3 m, REAL g, REAL r, REAL dt, REAL B0) {
    REAL th = -(g * r) - dt * (REAL)(2.0 * M_PI) * B0;
    REAL s = SIN(th), c = COS(th);
    V3 o;
    o.x = c * m.x - s * m.y;
    o.y = c * m.y + s * m.x;
    o.z = m.z;
    return o;
}

/* decay + regrowth: isochromat.cuh:99-107 */
static inline V3 SFX(iso_relax)(V3 m, REAL E1, REAL E2) {
    m.x *= E2;
    m.y *= E2;
    m.z = m.z * E1 + ((REAL)1 - E1);
    return m;
}

/*
 * simulate_pssfp_for_voxel: pssfp_kernels.cuh:14-108, summed over all z
 * positions (the reference's batches of 32/16/../1 only change the launch
 * layout, pssfp.cu:40-69).  echos: [nTR][nv] complex.
 */
void SFX(pssfp_simulate)(
    int nv, const float* T1, const float* T2, const float* B1, const float* B0,
    int nTR, const float* RF, int nRF, const float* gamma_dt_RF,
    float dt_ex, float dt_inv, float dt_pr, float g_ex, float g_inv, float g_pr,
    int nz, const float* zpos, REAL* echos, int nthreads) {
    long v;
    (void)g_inv;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4)
    for (v = 0; v < nv; v++) {
        REAL t1 = T1[v], t2 = T2[v], b1 = B1[v], b0 = B0[v];
        REAL E1_ex = SFX(relax)(dt_ex, t1), E2_ex = SFX(relax)(dt_ex, t2);
        REAL E1_pr = SFX(relax)(dt_pr, t1), E2_pr = SFX(relax)(dt_pr, t2);
        REAL E1_inv = SFX(relax)(dt_inv, t1), E2_inv = SFX(relax)(dt_inv, t2);
        int n, i, iz;
        for (n = 0; n < nTR; n++) C_ST(echos, (long)n * nv + v, C_MK(0, 0));
        for (iz = 0; iz < nz; iz++) {
            REAL z = zpos[iz];
            V3 m;
            CPX theta;
            m.x = 0;
            m.y = 0;
            m.z = 1;
            /* invert(p): isochromat.cuh:109-122 */
            m.z = COS((REAL)M_PI * b1) * m.z;
            m = SFX(iso_relax)(m, E1_inv, E2_inv);
            /* alpha/2 prepulse: pssfp_kernels.cuh:72-75 */
            theta = C_SCALE(C_LDC(RF, 0), (REAL)-0.5);
            for (i = 0; i < nRF; i++) {
                CPX zap = C_MUL(theta, C_LDC(gamma_dt_RF, i));
                m = SFX(iso_rotate_rf)(m, zap, g_ex, z, dt_ex, b1, b0);
                m = SFX(iso_relax)(m, E1_ex, E2_ex);
            }
            m = SFX(iso_rotate_z)(m, (REAL)2 * (REAL)g_pr, z, dt_pr, b0);
            m = SFX(iso_relax)(m, E1_pr, E2_pr);
            for (n = 0; n < nTR; n++) {
                CPX acc;
                theta = C_LDC(RF, n);
                for (i = 0; i < nRF; i++) {
                    CPX zap = C_MUL(theta, C_LDC(gamma_dt_RF, i));
                    m = SFX(iso_rotate_rf)(m, zap, g_ex, z, dt_ex, b1, b0);
                    m = SFX(iso_relax)(m, E1_ex, E2_ex);
                }
                m = SFX(iso_rotate_z)(m, g_pr, z, dt_pr, b0);
                m = SFX(iso_relax)(m, E1_pr, E2_pr);
                acc = C_LDR(echos, (long)n * nv + v);
                acc.re += m.x;
                acc.im += m.y;
                C_ST(echos, (long)n * nv + v, acc);
                m = SFX(iso_rotate_z)(m, g_pr, z, dt_pr, b0);
                m = SFX(iso_relax)(m, E1_pr, E2_pr);
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* k-space signal                                                            */
/* ------------------------------------------------------------------------ */

/*
 * Cartesian readout factor and per-sample exponent:
 * CartesianTrajectoryView::calculate_readout_magnetization (cartesian_view.cuh:28-44,
 * k_start unused, k0 = -ns/2 * delta_k) and calculate_sample_phase_decay (:52-63).
 * rho is applied as in prepare_readout_echos (signal_kernels.cuh:24-28).
 */
static inline CPX SFX(cart_factor)(int ns, REAL dt, REAL dk_re, REAL T2, REAL B0, CPX rho, REAL x) {
    REAL R2 = (REAL)1 / T2;
    CPX arg = C_MK(-R2, (REAL)(2.0 * M_PI) * B0);
    CPX res = C_SCALE(arg, (REAL)-0.5 * (REAL)ns * dt);
    REAL k0 = (REAL)-0.5 * (REAL)ns * dk_re;
    res.im += k0 * x;
    return C_MUL(C_EXP(res), rho);
}
static inline CPX SFX(cart_exponent)(REAL dt, REAL dk_re, REAL dk_im, REAL T2, REAL B0, REAL x, REAL y) {
    REAL R2 = (REAL)1 / T2;
    REAL Theta = dk_re * x + dk_im * y;
    Theta += dt * (REAL)(2.0 * M_PI) * B0;
    return C_MK(-dt * R2, Theta);
}

/*
 * magnetization_to_signal, Cartesian: S[c][r][s] = sum_v (A[r,v] * E_v^s) * C[c,v]
 * with A[r,v] = m[r,v] * factor_v (sum_signal_cartesian_naive,
 * signal_kernels.cuh:137-167; the Direct/GEMM methods of signal.cu:13-167 are
 * the same sum reordered).  E_v^s = exp(s * z_v) is evaluated directly.
 * Parallel over (coil, readout) rows.
 */
void SFX(signal_cartesian)(
    int nv, const float* T2, const float* B0, const float* rho, const float* x, const float* y,
    int nr, int ns, float dt, float dk_re, float dk_im,
    int nc, const float* coils, const REAL* echos, REAL* signal, int nthreads) {
    CPX* fac = (CPX*)malloc(sizeof(CPX) * (size_t)nv);
    CPX* Es = (CPX*)malloc(sizeof(CPX) * (size_t)nv * (size_t)ns);
    long v, row;
#pragma omp parallel for num_threads(nthreads)
    for (v = 0; v < nv; v++) {
        int s;
        CPX z = SFX(cart_exponent)(dt, dk_re, dk_im, T2[v], B0[v], x[v], y[v]);
        fac[v] = SFX(cart_factor)(ns, dt, dk_re, T2[v], B0[v], C_LDC(rho, v), x[v]);
        for (s = 0; s < ns; s++) Es[(size_t)s * nv + v] = C_EXP(C_SCALE(z, (REAL)s));
    }
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
    for (row = 0; row < (long)nc * nr; row++) {
        int c = (int)(row / nr), r = (int)(row % nr), s;
        const float* C = coils + 2L * c * nv;
        for (s = 0; s < ns; s++) {
            const CPX* E = Es + (size_t)s * nv;
            REAL are = 0, aim = 0;
            for (v = 0; v < nv; v++) {
                CPX a = C_MUL(C_LDR(echos, (long)r * nv + v), fac[v]);
                CPX t = C_MUL(C_MUL(a, E[v]), C_LDC(C, v));
                are += t.re;
                aim += t.im;
            }
            C_ST(signal, ((long)c * nr + r) * ns + s, C_MK(are, aim));
        }
    }
    free(fac);
    free(Es);
}

/*
 * magnetization_to_signal, spiral: signal.cu:169-240, prepare_readout_echos with
 * SpiralTrajectoryView::calculate_readout_magnetization (spiral_view.cuh:13-57:
 * rewind by ns/2*dt, prephaser uses only Re(k_start)*x), per-readout exponent
 * calculate_sample_phase_decay (:59-73) and sum_signal_spiral
 * (signal_kernels.cuh:274-345), with E^s evaluated as exp(s * z) (the
 * reference's pow(E, s0) * E^(s - s0) is the same quantity for integer s).
 */
void SFX(signal_spiral)(
    int nv, const float* T2, const float* B0, const float* rho, const float* x, const float* y,
    int nr, int ns, float dt, const float* k_start, const float* dk,
    int nc, const float* coils, const REAL* echos, REAL* signal, int nthreads) {
    long row;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
    for (row = 0; row < (long)nc * nr; row++) {
        int c = (int)(row / nr), r = (int)(row % nr), s;
        const float* C = coils + 2L * c * nv;
        CPX k0 = C_LDC(k_start, r), dkr = C_LDC(dk, r);
        long v;
        REAL* accre = (REAL*)calloc((size_t)ns, sizeof(REAL));
        REAL* accim = (REAL*)calloc((size_t)ns, sizeof(REAL));
        for (v = 0; v < nv; v++) {
            REAL R2 = (REAL)1 / (REAL)T2[v];
            CPX rw = C_EXP(C_SCALE(C_MK(R2, -(REAL)(2.0 * M_PI) * (REAL)B0[v]), (REAL)(0.5 * ns) * (REAL)dt));
            CPX pre = C_EXP(C_MK(0, k0.re * (REAL)x[v]));
            CPX a = C_MUL(C_MUL(C_MUL(C_LDR(echos, (long)r * nv + v), rw), pre), C_LDC(rho, v));
            REAL Theta = dkr.re * (REAL)x[v] + dkr.im * (REAL)y[v];
            CPX z, cv = C_LDC(C, v);
            Theta += (REAL)dt * (REAL)(2.0 * M_PI) * (REAL)B0[v];
            z = C_MK(-(REAL)dt * R2, Theta);
            for (s = 0; s < ns; s++) {
                CPX t = C_MUL(C_MUL(a, C_EXP(C_SCALE(z, (REAL)s))), cv);
                accre[s] += t.re;
                accim[s] += t.im;
            }
        }
        for (s = 0; s < ns; s++) C_ST(signal, ((long)c * nr + r) * ns + s, C_MK(accre[s], accim[s]));
        free(accre);
        free(accim);
    }
}

/* phase_encoding: phase_encoding_kernels.cuh:10-25 -- m * exp(i * y * Im k_start[r]) */
void SFX(phase_encoding)(int nv, const float* y, int nr, const float* k_start, const REAL* echos, REAL* out, int nthreads) {
    long r;
#pragma omp parallel for num_threads(nthreads)
    for (r = 0; r < nr; r++) {
        REAL ky = (REAL)k_start[2 * r + 1];
        long v;
        for (v = 0; v < nv; v++) {
            CPX e = C_EXP(C_MK(0, (REAL)y[v] * ky));
            C_ST(out, r * nv + v, C_MUL(C_LDR(echos, r * nv + v), e));
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Jacobian products (Cartesian, 4 parameters: T1, T2, rho_x, rho_y)          */
/* ------------------------------------------------------------------------ */

/* calculate_sample_decay_absolute{,_delta_T2}: cartesian_view.cuh:65-92 */
static inline void SFX(abs_decay)(int s, int ns, REAL dt, REAL dk_re, REAL dk_im, REAL T2, REAL B0, REAL x, REAL y, CPX* E, CPX* dE) {
    REAL R2 = (REAL)1 / T2;
    REAL sp = (REAL)s - (REAL)0.5 * (REAL)ns;
    REAL Theta = dk_re * x + dk_im * y;
    Theta += dt * (REAL)(2.0 * M_PI) * B0;
    *E = C_EXP(C_SCALE(C_MK(-dt * R2, Theta), sp));
    *dE = C_SCALE(*E, sp * dt * R2 * R2);
}

/*
 * compute_jacobian: jacobian_product_naive (product_kernels.cuh:7-67).
 * v: [4][nv] complex; Jv: [nc][nr][ns] complex.
 */
void SFX(jacobian_product)(
    int nv, const float* T1, const float* T2, const float* B0, const float* rho, const float* x, const float* y,
    int nr, int ns, float dt, float dk_re, float dk_im, int nc, const float* coils,
    const REAL* echos, const REAL* dT1, const REAL* dT2, const REAL* vec, REAL* Jv, int nthreads) {
    CPX* E = (CPX*)malloc(sizeof(CPX) * (size_t)ns * nv);
    CPX* dE = (CPX*)malloc(sizeof(CPX) * (size_t)ns * nv);
    long row;
#pragma omp parallel for num_threads(nthreads)
    for (row = 0; row < (long)ns * nv; row++) {
        int s = (int)(row / nv);
        long v = row % nv;
        SFX(abs_decay)(s, ns, dt, dk_re, dk_im, T2[v], B0[v], x[v], y[v], &E[row], &dE[row]);
    }
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
    for (row = 0; row < (long)nc * nr; row++) {
        int c = (int)(row / nr), r = (int)(row % nr), s;
        for (s = 0; s < ns; s++) {
            REAL are = 0, aim = 0;
            long v;
            for (v = 0; v < nv; v++) {
                CPX Es = E[(size_t)s * nv + v], dEs = dE[(size_t)s * nv + v];
                CPX me = C_LDR(echos, (long)r * nv + v);
                CPX d1 = C_LDR(dT1, (long)r * nv + v), d2 = C_LDR(dT2, (long)r * nv + v);
                CPX rh = C_LDC(rho, v);
                CPX t0 = C_MUL(C_MUL(C_SCALE(C_MUL(C_LDR(vec, v), Es), (REAL)T1[v]), rh), d1);
                CPX t1 = C_MUL(C_SCALE(C_MUL(C_LDR(vec, (long)nv + v), C_ADD(C_MUL(Es, d2), C_MUL(dEs, me))), (REAL)T2[v]), rh);
                CPX t2 = C_MUL(C_MUL(C_LDR(vec, 2L * nv + v), Es), me);
                CPX t3 = C_MUL(C_MUL(C_MUL(C_LDR(vec, 3L * nv + v), Es), C_MK(0, 1)), me);
                CPX prod = C_ADD(C_ADD(t0, t1), C_ADD(t2, t3));
                CPX t = C_MUL(prod, C_LDC(coils, (long)c * nv + v));
                are += t.re;
                aim += t.im;
            }
            C_ST(Jv, ((long)c * nr + r) * ns + s, C_MK(are, aim));
        }
    }
    free(E);
    free(dE);
}

/*
 * compute_jacobian_hermitian: jacobian_hermitian_product_naive
 * (hermitian_kernels.cuh:30-92).  yv: [nc][nr][ns] complex; JHy: [4][nv].
 */
void SFX(jacobian_hermitian)(
    int nv, const float* T1, const float* T2, const float* B0, const float* rho, const float* x, const float* y,
    int nr, int ns, float dt, float dk_re, float dk_im, int nc, const float* coils,
    const REAL* echos, const REAL* dT1, const REAL* dT2, const REAL* yv, REAL* JHy, int nthreads) {
    long v;
#pragma omp parallel num_threads(nthreads)
    {
        CPX* E = (CPX*)malloc(sizeof(CPX) * (size_t)ns);
        CPX* dE = (CPX*)malloc(sizeof(CPX) * (size_t)ns);
#pragma omp for schedule(dynamic, 8)
        for (v = 0; v < nv; v++) {
            CPX out[4];
            int c, r, s, p;
            CPX rh = C_LDC(rho, v);
            for (p = 0; p < 4; p++) out[p] = C_MK(0, 0);
            for (s = 0; s < ns; s++) SFX(abs_decay)(s, ns, dt, dk_re, dk_im, T2[v], B0[v], x[v], y[v], &E[s], &dE[s]);
            for (c = 0; c < nc; c++) {
                CPX mHv = C_MK(0, 0), dm0 = C_MK(0, 0), dm1 = C_MK(0, 0);
                CPX cc = C_LDC(coils, (long)c * nv + v);
                for (r = 0; r < nr; r++) {
                    CPX me = C_LDR(echos, (long)r * nv + v);
                    CPX d1 = C_LDR(dT1, (long)r * nv + v), d2 = C_LDR(dT2, (long)r * nv + v);
                    for (s = 0; s < ns; s++) {
                        CPX f = C_LDR(yv, ((long)c * nr + r) * ns + s);
                        CPX ms = C_MUL(E[s], me);
                        CPX dms0 = C_MUL(d1, E[s]);
                        CPX dms1 = C_ADD(C_MUL(d2, E[s]), C_MUL(me, dE[s]));
                        mHv = C_ADD(mHv, C_MUL(f, C_CONJ(ms)));
                        dm0 = C_ADD(dm0, C_MUL(f, C_CONJ(dms0)));
                        dm1 = C_ADD(dm1, C_MUL(f, C_CONJ(dms1)));
                    }
                }
                out[0] = C_ADD(out[0], C_MUL(C_MUL(C_CONJ(C_SCALE(rh, (REAL)T1[v])), C_CONJ(cc)), dm0));
                out[1] = C_ADD(out[1], C_MUL(C_MUL(C_CONJ(C_SCALE(rh, (REAL)T2[v])), C_CONJ(cc)), dm1));
                out[2] = C_ADD(out[2], C_MUL(C_CONJ(cc), mHv));
                out[3] = C_ADD(out[3], C_MUL(C_MUL(C_CONJ(cc), mHv), C_MK(0, -1)));
            }
            for (p = 0; p < 4; p++) C_ST(JHy, (long)p * nv + v, out[p]);
        }
        free(E);
        free(dE);
    }
}

/* ------------------------------------------------------------------------ */
/* Jacobian products on the spiral trajectory, 6 parameters                 */
/* (T1, T2, B1, B0, rho_x, rho_y).  The reference has no spiral Jacobian:    */
/* these are the exact derivatives of its spiral signal (the S of            */
/* signal_spiral above, spiral_view.cuh:13-73) with the parameter scaling of */
/* its Cartesian J.v (product_kernels.cuh:56-60: the T1 and T2 rows are      */
/* scaled by T1 and T2, the rho rows multiply m and i m).  With              */
/*   Phi[r][s][v] = exp((s - ns/2) dt (-R2 + i 2 pi B0)                      */
/*                      + i (Re k0_r x + s (Re dk_r x + Im dk_r y)))         */
/* S = sum_v C rho m Phi, dPhi/dT2 = (s - ns/2) dt R2^2 Phi and              */
/* dPhi/dB0 = (s - ns/2) i 2 pi dt Phi.  Pinned by finite differences of     */
/* signal_spiral and by the adjoint identity <y, J v> = <J^H y, v>.          */
/* ------------------------------------------------------------------------ */

/* Phi[r][0][v] (readout factor of signal_spiral) and the per-sample exponent */
static inline void SFX(spiral_phi)(int ns, REAL dt, CPX k0, CPX dkr, REAL T2, REAL B0, REAL x, REAL y, CPX* phi0, CPX* z) {
    REAL R2 = (REAL)1 / T2;
    CPX rw = C_EXP(C_SCALE(C_MK(R2, -(REAL)(2.0 * M_PI) * B0), (REAL)(0.5 * ns) * dt));
    CPX pre = C_EXP(C_MK(0, k0.re * x));
    REAL Theta = dkr.re * x + dkr.im * y;
    *phi0 = C_MUL(rw, pre);
    Theta += dt * (REAL)(2.0 * M_PI) * B0;
    *z = C_MK(-dt * R2, Theta);
}

/*
 * J.v, spiral, 6 parameters.  vec: [6][nv] complex (T1, T2, B1, B0, rho_x,
 * rho_y); d*: [nr][nv] derivatives of the magnetisation; Jv: [nc][nr][ns].
 *   Jv = sum_v C (a + (s - ns/2) b) Phi,
 *   a = rho (v0 T1 dm_T1 + v1 T2 dm_T2 + v2 dm_B1 + v3 dm_B0) + (v4 + i v5) m
 *   b = rho m dt (v1 T2 R2^2 + v3 i 2 pi)
 */
void SFX(jacobian_product_spiral)(
    int nv, const float* T1, const float* T2, const float* B0, const float* rho, const float* x, const float* y,
    int nr, int ns, float dt, const float* k_start, const float* dk, int nc, const float* coils,
    const REAL* echos, const REAL* dT1, const REAL* dT2, const REAL* dB1, const REAL* dB0, const REAL* vec,
    REAL* Jv, int nthreads) {
    long row;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
    for (row = 0; row < (long)nc * nr; row++) {
        int c = (int)(row / nr), r = (int)(row % nr), s;
        CPX k0 = C_LDC(k_start, r), dkr = C_LDC(dk, r);
        long v;
        REAL* accre = (REAL*)calloc((size_t)ns, sizeof(REAL));
        REAL* accim = (REAL*)calloc((size_t)ns, sizeof(REAL));
        for (v = 0; v < nv; v++) {
            long i = (long)r * nv + v;
            REAL R2 = (REAL)1 / (REAL)T2[v];
            CPX phi0, z, a, b, cv = C_LDC(coils, (long)c * nv + v), rh = C_LDC(rho, v), me = C_LDR(echos, i);
            SFX(spiral_phi)(ns, (REAL)dt, k0, dkr, (REAL)T2[v], (REAL)B0[v], (REAL)x[v], (REAL)y[v], &phi0, &z);
            a = C_SCALE(C_MUL(C_LDR(vec, v), C_LDR(dT1, i)), (REAL)T1[v]);
            a = C_ADD(a, C_SCALE(C_MUL(C_LDR(vec, (long)nv + v), C_LDR(dT2, i)), (REAL)T2[v]));
            a = C_ADD(a, C_MUL(C_LDR(vec, 2L * nv + v), C_LDR(dB1, i)));
            a = C_ADD(a, C_MUL(C_LDR(vec, 3L * nv + v), C_LDR(dB0, i)));
            a = C_MUL(a, rh);
            a = C_ADD(a, C_MUL(C_ADD(C_LDR(vec, 4L * nv + v), C_MUL(C_LDR(vec, 5L * nv + v), C_MK(0, 1))), me));
            b = C_ADD(C_SCALE(C_LDR(vec, (long)nv + v), (REAL)T2[v] * R2 * R2),
                      C_MUL(C_LDR(vec, 3L * nv + v), C_MK(0, (REAL)(2.0 * M_PI))));
            b = C_SCALE(C_MUL(C_MUL(b, rh), me), (REAL)dt);
            for (s = 0; s < ns; s++) {
                CPX phi = C_MUL(phi0, C_EXP(C_SCALE(z, (REAL)s)));
                CPX amp = C_ADD(a, C_SCALE(b, (REAL)s - (REAL)0.5 * (REAL)ns));
                CPX t = C_MUL(C_MUL(amp, phi), cv);
                accre[s] += t.re;
                accim[s] += t.im;
            }
        }
        for (s = 0; s < ns; s++) C_ST(Jv, ((long)c * nr + r) * ns + s, C_MK(accre[s], accim[s]));
        free(accre);
        free(accim);
    }
}

/*
 * J^H.y, spiral, 6 parameters: the adjoint of jacobian_product_spiral.
 * yv: [nc][nr][ns]; JHy: [6][nv].  Per voxel, coil and readout
 *   U = sum_s y conj(Phi),  U' = sum_s (s - ns/2) y conj(Phi)
 * and JHy = sum_c (conj(T1 rho C) sum_r conj(dm_T1) U,
 *                  conj(T2 rho C) sum_r [conj(dm_T2) U + conj(m) dt R2^2 U'],
 *                  conj(rho C) sum_r conj(dm_B1) U,
 *                  conj(rho C) sum_r [conj(dm_B0) U - i 2 pi dt conj(m) U'],
 *                  conj(C) sum_r conj(m) U,  -i conj(C) sum_r conj(m) U).
 */
void SFX(jacobian_hermitian_spiral)(
    int nv, const float* T1, const float* T2, const float* B0, const float* rho, const float* x, const float* y,
    int nr, int ns, float dt, const float* k_start, const float* dk, int nc, const float* coils,
    const REAL* echos, const REAL* dT1, const REAL* dT2, const REAL* dB1, const REAL* dB0, const REAL* yv,
    REAL* JHy, int nthreads) {
    long v;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 8)
    for (v = 0; v < nv; v++) {
        CPX out[6];
        int c, r, s, p;
        REAL R2 = (REAL)1 / (REAL)T2[v];
        CPX rh = C_LDC(rho, v);
        for (p = 0; p < 6; p++) out[p] = C_MK(0, 0);
        for (c = 0; c < nc; c++) {
            CPX g1 = C_MK(0, 0), g2 = C_MK(0, 0), gb1 = C_MK(0, 0), gb0 = C_MK(0, 0), u = C_MK(0, 0);
            CPX cc = C_LDC(coils, (long)c * nv + v);
            for (r = 0; r < nr; r++) {
                long i = (long)r * nv + v;
                CPX phi0, z, U = C_MK(0, 0), Up = C_MK(0, 0);
                CPX mc = C_CONJ(C_LDR(echos, i));
                SFX(spiral_phi)(ns, (REAL)dt, C_LDC(k_start, r), C_LDC(dk, r), (REAL)T2[v], (REAL)B0[v], (REAL)x[v],
                                (REAL)y[v], &phi0, &z);
                for (s = 0; s < ns; s++) {
                    CPX phi = C_MUL(phi0, C_EXP(C_SCALE(z, (REAL)s)));
                    CPX t = C_MUL(C_LDR(yv, ((long)c * nr + r) * ns + s), C_CONJ(phi));
                    U = C_ADD(U, t);
                    Up = C_ADD(Up, C_SCALE(t, (REAL)s - (REAL)0.5 * (REAL)ns));
                }
                g1 = C_ADD(g1, C_MUL(C_CONJ(C_LDR(dT1, i)), U));
                g2 = C_ADD(g2, C_ADD(C_MUL(C_CONJ(C_LDR(dT2, i)), U), C_SCALE(C_MUL(mc, Up), (REAL)dt * R2 * R2)));
                gb1 = C_ADD(gb1, C_MUL(C_CONJ(C_LDR(dB1, i)), U));
                gb0 = C_ADD(gb0, C_ADD(C_MUL(C_CONJ(C_LDR(dB0, i)), U),
                                       C_MUL(C_MUL(mc, Up), C_MK(0, -(REAL)(2.0 * M_PI) * (REAL)dt))));
                u = C_ADD(u, C_MUL(mc, U));
            }
            out[0] = C_ADD(out[0], C_MUL(C_CONJ(C_MUL(C_SCALE(rh, (REAL)T1[v]), cc)), g1));
            out[1] = C_ADD(out[1], C_MUL(C_CONJ(C_MUL(C_SCALE(rh, (REAL)T2[v]), cc)), g2));
            out[2] = C_ADD(out[2], C_MUL(C_CONJ(C_MUL(rh, cc)), gb1));
            out[3] = C_ADD(out[3], C_MUL(C_CONJ(C_MUL(rh, cc)), gb0));
            out[4] = C_ADD(out[4], C_MUL(C_CONJ(cc), u));
            out[5] = C_ADD(out[5], C_MUL(C_MUL(C_CONJ(cc), u), C_MK(0, -1)));
        }
        for (p = 0; p < 6; p++) C_ST(JHy, (long)p * nv + v, out[p]);
    }
}

#undef C_MK
#undef C_ADD
#undef C_SUB
#undef C_MUL
#undef C_SCALE
#undef C_CONJ
#undef C_EXP
#undef C_POLAR
#undef C_LDC
#undef C_LDR
#undef C_ST
#undef CPX
#undef V3
