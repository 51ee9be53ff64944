/*
 * CPU oracle for the staggered-grid FD time step: the type-generic body.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is part of the oracle: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker.  The product (rockseis_amd, librsgpu.so) never links
 * or calls it.
 *
 * rsoracle.c includes this file twice: with REAL=float (suffix _f32, the
 * reference's own T=float path, used for parity) and REAL=double (suffix
 * _f64, the FP64 oracle used for error statements).  Every arithmetic
 * expression keeps the reference's operand order and its float/double
 * promotions so the f32 build reproduces the reference bit for bit when both
 * are compiled without FMA contraction (-ffp-contract=off, no -march).
 *
 * Layout (reference waves.cpp:4-13): x fastest, I2D(i,j)=j*nx+i,
 * I3D(i,j,k)=k*nx*ny+j*nx+i.  PML strips: left/right I2D_lr=j*lpml+i,
 * top/bottom I2D_tb=j*nx+i; 3D I3D_lr=k*lpml*ny+j*lpml+i,
 * I3D_fb=k*nx*lpml+j*nx+i, I3D_tb=k*nx*ny+j*nx+i.
 *
 * Domain decomposition (Domain::getStatus()==true) is not restated: all
 * functions follow the single-domain branch (ix0=iy0=iz0=0, all six PML
 * sides applied).
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SUF)

/* ---------------------------------------------------------------------- */
/* Der: staggered first derivatives (der.cpp:155-540)                      */
/* ---------------------------------------------------------------------- */

/* Coefficient table as stored by Der<T>'s constructor (der.cpp:77-136).
 * w11 is a float literal in the reference, the others double literals;
 * CARRAYSIZE=36, the set for stencil order k starts at cind(k-1)=k(k-1)/2. */
static REAL FN(coef_)[36];
static int FN(coef_ready_) = 0;

static const REAL *FN(coefs)(void)
{
    if (!FN(coef_ready_)) {
        for (int i = 0; i < 36; i++) FN(coef_)[i] = (REAL)orc_der_table[i];
        FN(coef_)[0] = (REAL)orc_w11;
        FN(coef_ready_) = 1;
    }
    return FN(coef_);
}

void FN(orc_der_coeffs)(REAL *out36)
{
    const REAL *c = FN(coefs)();
    for (int i = 0; i < 36; i++) out36[i] = c[i];
}

/*
 * One derivative pass along `axis` (0=x, 1=y, 2=z), forward (fw=1,
 * der.cpp:155-346) or backward (fw=0, der.cpp:351-540).  At index i of an
 * n-long axis the reference uses the stencil order
 *     forward : k = min(order, i+1, n-1-i)   (k=0 -> df=0, i.e. i=n-1)
 *     backward: k = min(order, i,   n-i)     (k=0 -> df=0, i.e. i=0)
 * which is what its left/interior/right loops evaluate for n >= 2*order.
 * Accumulation order: term 0, then io=1..k-1, then one division by d.
 */
void FN(orc_deriv)(const REAL *f, REAL *df, int nx, int ny, int nz,
                   int axis, int fw, int order, REAL d)
{
    const REAL *coef = FN(coefs)();
    long sxy = (long)nx * ny;
    long s = axis == 0 ? 1 : (axis == 1 ? nx : sxy);
    int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
    if (order < 1) order = 1;
    if (order > 8) order = 8;
    for (int iz = 0; iz < nz; iz++)
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                long p = iz * sxy + (long)iy * nx + ix;
                int i = axis == 0 ? ix : (axis == 1 ? iy : iz);
                int k = fw ? orc_kfw(i, n, order) : orc_kbw(i, n, order);
                if (k == 0) {
                    df[p] = 0.0;
                    continue;
                }
                const REAL *c = coef + ((k - 1) * k) / 2;
                REAL acc;
                if (fw) {
                    acc = (f[p + s] - f[p]) * c[0];
                    for (int io = 1; io < k; io++)
                        acc += (f[p + (io + 1) * s] - f[p - io * s]) * c[io];
                } else {
                    acc = (f[p] - f[p - s]) * c[0];
                    for (int io = 1; io < k; io++)
                        acc += (f[p + io * s] - f[p - (io + 1) * s]) * c[io];
                }
                acc /= d;
                df[p] = acc;
            }
}

/* ---------------------------------------------------------------------- */
/* Pml::computeABC (pml.cpp:79-126)                                        */
/* ---------------------------------------------------------------------- */

/* One profile point: func^3 scaled into a,k,s then C,B,A.  The reference
 * evaluates pow() in double (C++11 std::pow(float,int) promotes), exp() in
 * T, and the C and A quotients in double before storing into T. */
static void FN(abc_point_)(REAL func, REAL Amax, REAL Kmax, REAL Smax, REAL dt,
                           REAL *A, REAL *B, REAL *C)
{
    REAL a, k, s;
    func = (REAL)pow((double)func, 3.0);
    a = Amax * func;
    k = (REAL)(1. + (double)((Kmax - 1) * func));
    s = Smax * func;
    *C = (REAL)(1.0 - 1.0 / (double)k);
    *B = REXP(-(a + s / k) * dt);
    *A = (REAL)(((double)s * (1. - (double)*B)) / (double)(k * (k * a + s)));
}

/* tab is 12*lpml: A_ltf,B_ltf,C_ltf, A_ltf_stag,B_ltf_stag,C_ltf_stag,
 * A_rbb,B_rbb,C_rbb, A_rbb_stag,B_rbb_stag,C_rbb_stag (pml.h:60-74). */
void FN(orc_pml_abc)(int lpml, REAL dt, REAL Amax, REAL Kmax, REAL Smax,
                     int sign, REAL *tab)
{
    REAL *t[12];
    for (int j = 0; j < 12; j++) t[j] = tab + (long)j * lpml;
    for (int i = 0; i < lpml; i++) {
        REAL func;
        func = (REAL)(lpml - i) / lpml;
        FN(abc_point_)(func, Amax, Kmax, Smax, dt, &t[0][i], &t[1][i], &t[2][i]);
        func = (REAL)((double)(lpml - i) - sign * 0.5) / lpml;
        FN(abc_point_)(func, Amax, Kmax, Smax, dt, &t[3][i], &t[4][i], &t[5][i]);
        func = (REAL)((double)i + 1.0) / lpml;
        FN(abc_point_)(func, Amax, Kmax, Smax, dt, &t[6][i], &t[7][i], &t[8][i]);
        func = (REAL)((double)i + 1.0 + sign * 0.5) / lpml;
        FN(abc_point_)(func, Amax, Kmax, Smax, dt, &t[9][i], &t[10][i], &t[11][i]);
    }
}

/* ---------------------------------------------------------------------- */
/* Shared PML strip update: psi = B*psi + A*df; F -= dt*M*(psi + C*df)     */
/* ---------------------------------------------------------------------- */

#define PSI_UPD(psi, Bt, At, Ct, dfv, F, Mv, dt)                              \
    do {                                                                    \
        (psi) = (Bt) * (psi) + (At) * (dfv);                                \
        (F) -= (dt) * (Mv) * ((psi) + (Ct) * (dfv));                        \
    } while (0)
/* Same with no modulus (WavesElastic2D::backwardstepStress) */
#define PSI_UPD0(psi, Bt, At, Ct, dfv, F, dt)                                 \
    do {                                                                    \
        (psi) = (Bt) * (psi) + (At) * (dfv);                                \
        (F) -= (dt) * ((psi) + (Ct) * (dfv));                               \
    } while (0)

/* Table pointers: stag=1 picks the staggered set */
#define TAB_A_L(tab, L, stag) ((tab) + (long)((stag) ? 3 : 0) * (L))
#define TAB_B_L(tab, L, stag) ((tab) + (long)((stag) ? 4 : 1) * (L))
#define TAB_C_L(tab, L, stag) ((tab) + (long)((stag) ? 5 : 2) * (L))
#define TAB_A_R(tab, L, stag) ((tab) + (long)((stag) ? 9 : 6) * (L))
#define TAB_B_R(tab, L, stag) ((tab) + (long)((stag) ? 10 : 7) * (L))
#define TAB_C_R(tab, L, stag) ((tab) + (long)((stag) ? 11 : 8) * (L))

/*
 * Add dt*M*df over the full 2D grid then apply the x-sides (left/right)
 * strips (waves.cpp:225-267 pattern).  M==NULL means modulus 1 (backward
 * stress steps).  F2/M2 optional second target sharing the same psi (the
 * elastic Sxx/Szz pair).
 */
static void FN(upd_x2d_)(int nx, int nz, int L, REAL dt, const REAL *tab, int stag,
                         const REAL *df, REAL *F, const REAL *M, REAL *F2,
                         const REAL *M2, REAL *psiL, REAL *psiR)
{
    long n = (long)nx * nz;
    for (long p = 0; p < n; p++) {
        if (M) F[p] += dt * M[p] * df[p]; else F[p] += dt * df[p];
        if (F2) F2[p] += dt * M2[p] * df[p];
    }
    const REAL *Al = TAB_A_L(tab, L, stag), *Bl = TAB_B_L(tab, L, stag), *Cl = TAB_C_L(tab, L, stag);
    const REAL *Ar = TAB_A_R(tab, L, stag), *Br = TAB_B_R(tab, L, stag), *Cr = TAB_C_R(tab, L, stag);
    for (int iz = 0; iz < nz; iz++)
        for (int ix = 0; ix < L; ix++) {
            long p = (long)iz * nx + ix;
            long q = (long)iz * L + ix;
            if (M) PSI_UPD(psiL[q], Bl[ix], Al[ix], Cl[ix], df[p], F[p], M[p], dt);
            else PSI_UPD0(psiL[q], Bl[ix], Al[ix], Cl[ix], df[p], F[p], dt);
            if (F2) F2[p] -= dt * M2[p] * (psiL[q] + Cl[ix] * df[p]);
            long pr = (long)iz * nx + ix + nx - L;
            if (M) PSI_UPD(psiR[q], Br[ix], Ar[ix], Cr[ix], df[pr], F[pr], M[pr], dt);
            else PSI_UPD0(psiR[q], Br[ix], Ar[ix], Cr[ix], df[pr], F[pr], dt);
            if (F2) F2[pr] -= dt * M2[pr] * (psiR[q] + Cr[ix] * df[pr]);
        }
}

/* Same for the z-sides (top/bottom) strips (waves.cpp:279-307 pattern). */
static void FN(upd_z2d_)(int nx, int nz, int L, REAL dt, const REAL *tab, int stag,
                         const REAL *df, REAL *F, const REAL *M, REAL *F2,
                         const REAL *M2, REAL *psiT, REAL *psiB)
{
    long n = (long)nx * nz;
    for (long p = 0; p < n; p++) {
        if (M) F[p] += dt * M[p] * df[p]; else F[p] += dt * df[p];
        if (F2) F2[p] += dt * M2[p] * df[p];
    }
    const REAL *Al = TAB_A_L(tab, L, stag), *Bl = TAB_B_L(tab, L, stag), *Cl = TAB_C_L(tab, L, stag);
    const REAL *Ar = TAB_A_R(tab, L, stag), *Br = TAB_B_R(tab, L, stag), *Cr = TAB_C_R(tab, L, stag);
    for (int iz = 0; iz < L; iz++)
        for (int ix = 0; ix < nx; ix++) {
            long p = (long)iz * nx + ix;
            long q = p;
            if (M) PSI_UPD(psiT[q], Bl[iz], Al[iz], Cl[iz], df[p], F[p], M[p], dt);
            else PSI_UPD0(psiT[q], Bl[iz], Al[iz], Cl[iz], df[p], F[p], dt);
            if (F2) F2[p] -= dt * M2[p] * (psiT[q] + Cl[iz] * df[p]);
            long pb = (long)(iz + nz - L) * nx + ix;
            if (M) PSI_UPD(psiB[q], Br[iz], Ar[iz], Cr[iz], df[pb], F[pb], M[pb], dt);
            else PSI_UPD0(psiB[q], Br[iz], Ar[iz], Cr[iz], df[pb], F[pb], dt);
            if (F2) F2[pb] -= dt * M2[pb] * (psiB[q] + Cr[iz] * df[pb]);
        }
}

/* ---------------------------------------------------------------------- */
/* WavesAcoustic2D (waves.cpp:217-595)                                     */
/* ---------------------------------------------------------------------- */

typedef struct {
    int nx, nz, lpml, order; /* padded sizes */
    REAL dx, dz, dt;
    const REAL *tab;         /* 12*lpml */
    REAL *P, *Vx, *Vz;
    /* strips: P_left,P_right (lpml*nz) P_top,P_bottom (nx*lpml)
     *         Vxx_left,Vxx_right, Vzz_top,Vzz_bottom */
    REAL *P_left, *P_right, *P_top, *P_bottom;
    REAL *Vxx_left, *Vxx_right, *Vzz_top, *Vzz_bottom;
    const REAL *L, *Rx, *Rz;
    REAL *df; /* scratch nx*nz */
} FN(orc_a2d_t);

void FN(orc_a2d_velocity)(FN(orc_a2d_t) *s)
{
    FN(orc_deriv)(s->P, s->df, s->nx, 1, s->nz, 0, 1, s->order, s->dx);
    FN(upd_x2d_)(s->nx, s->nz, s->lpml, s->dt, s->tab, 1, s->df, s->Vx, s->Rx, 0, 0, s->P_left, s->P_right);
    FN(orc_deriv)(s->P, s->df, s->nx, 1, s->nz, 2, 1, s->order, s->dz);
    FN(upd_z2d_)(s->nx, s->nz, s->lpml, s->dt, s->tab, 1, s->df, s->Vz, s->Rz, 0, 0, s->P_top, s->P_bottom);
}

void FN(orc_a2d_stress)(FN(orc_a2d_t) *s)
{
    FN(orc_deriv)(s->Vx, s->df, s->nx, 1, s->nz, 0, 0, s->order, s->dx);
    FN(upd_x2d_)(s->nx, s->nz, s->lpml, s->dt, s->tab, 0, s->df, s->P, s->L, 0, 0, s->Vxx_left, s->Vxx_right);
    FN(orc_deriv)(s->Vz, s->df, s->nx, 1, s->nz, 2, 0, s->order, s->dz);
    FN(upd_z2d_)(s->nx, s->nz, s->lpml, s->dt, s->tab, 0, s->df, s->P, s->L, 0, 0, s->Vzz_top, s->Vzz_bottom);
}

/* ---------------------------------------------------------------------- */
/* Exact discrete adjoint of the acoustic 2D step (no reference            */
/* counterpart: the reference's FwiAcoustic2D back-propagates with the     */
/* forward operator).  Written from the forward restatement above: every   */
/* linear map of the step is transposed as it stands, edge stencils and    */
/* PML strips included, so <F x, y> = <x, F^T y> holds to rounding.        */
/* ---------------------------------------------------------------------- */

/* out[j] += sum_i D[i,j] u[i] for the derivative matrix D of orc_deriv
 * (same axis, direction, order and step): row i of D holds +c/d at the
 * cells orc_deriv adds and -c/d at those it subtracts. */
void FN(orc_deriv_t)(const REAL *u, REAL *out, int nx, int ny, int nz,
                     int axis, int fw, int order, REAL d)
{
    const REAL *coef = FN(coefs)();
    long sxy = (long)nx * ny;
    long s = axis == 0 ? 1 : (axis == 1 ? nx : sxy);
    int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
    if (order < 1) order = 1;
    if (order > 8) order = 8;
    for (int iz = 0; iz < nz; iz++)
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                long p = iz * sxy + (long)iy * nx + ix;
                int i = axis == 0 ? ix : (axis == 1 ? iy : iz);
                int k = fw ? orc_kfw(i, n, order) : orc_kbw(i, n, order);
                if (k == 0) continue;
                const REAL *c = coef + ((k - 1) * k) / 2;
                REAL ui = u[p] / d;
                for (int io = 0; io < k; io++) {
                    if (fw) {
                        out[p + (io + 1) * s] += c[io] * ui;
                        out[p - io * s] -= c[io] * ui;
                    } else {
                        out[p + io * s] += c[io] * ui;
                        out[p - (io + 1) * s] -= c[io] * ui;
                    }
                }
            }
}

/*
 * Reverse of one derivative term of upd_x2d_ (axis 0) / upd_z2d_ (axis 2).
 * Forward, with d the derivative, M the modulus and, in a strip cell,
 * psi' = B*psi + A*d:
 *     F' = F + dt*M*d                          (outside the strips)
 *     F' = F + dt*M*(1-A-C)*d - dt*M*B*psi     (inside)
 * Given lF = adjoint of F' and lpsi = adjoint of psi', this writes
 * w = adjoint of d, replaces lpsi by the adjoint of the old psi and adds the
 * gradient of M, lF * dF'/dM, to g.  psi holds the forward strip before
 * (psi_old=1) or after (psi_old=0) the update.
 */
static void FN(adj_upd_2d_)(int axis, int nx, int nz, int L, REAL dt, const REAL *tab, int stag,
                            const REAL *df, const REAL *lF, const REAL *M, const REAL *psiLo,
                            const REAL *psiHi, int psi_old, REAL *lpsiLo, REAL *lpsiHi, REAL *w,
                            REAL *g)
{
    long n = (long)nx * nz;
    for (long p = 0; p < n; p++) {
        w[p] = dt * M[p] * lF[p];
        g[p] += lF[p] * dt * df[p];
    }
    int n1 = axis == 0 ? nz : L, n2 = axis == 0 ? L : nx;
    for (int side = 0; side < 2; side++) {
        const REAL *At = side ? TAB_A_R(tab, L, stag) : TAB_A_L(tab, L, stag);
        const REAL *Bt = side ? TAB_B_R(tab, L, stag) : TAB_B_L(tab, L, stag);
        const REAL *Ct = side ? TAB_C_R(tab, L, stag) : TAB_C_L(tab, L, stag);
        const REAL *psi = side ? psiHi : psiLo;
        REAL *lpsi = side ? lpsiHi : lpsiLo;
        for (int i1 = 0; i1 < n1; i1++)
            for (int i2 = 0; i2 < n2; i2++) {
                long q = (long)i1 * n2 + i2, p;
                int t;
                if (axis == 0) {
                    t = i2;
                    p = (long)i1 * nx + i2 + (side ? nx - L : 0);
                } else {
                    t = i1;
                    p = (long)(i1 + (side ? nz - L : 0)) * nx + i2;
                }
                REAL A = At[t], B = Bt[t], C = Ct[t];
                REAL pn = psi_old ? B * psi[q] + A * df[p] : psi[q];
                REAL lf = lF[p], lq = lpsi[q], dtM = dt * M[p];
                w[p] = dtM * (1 - A - C) * lf + A * lq;
                lpsi[q] = B * lq - dtM * B * lf;
                g[p] -= lf * dt * (pn + C * df[p]);
            }
    }
}

/* Reverse of the stress half step.  s: forward state holding V of this
 * step and the Vxx/Vzz strips before it (the forward velocity half step
 * applied, the stress half step not); a: adjoint state (P = adjoint of the
 * pressure after the step, unchanged; Vx, Vz and the Vxx/Vzz strips are
 * carried back through the step; a->df is scratch).  gL += dJ/dL. */
void FN(orc_a2d_adj_stress)(FN(orc_a2d_t) *s, FN(orc_a2d_t) *a, REAL *gL)
{
    int nx = s->nx, nz = s->nz;
    FN(orc_deriv)(s->Vx, s->df, nx, 1, nz, 0, 0, s->order, s->dx);
    FN(adj_upd_2d_)(0, nx, nz, s->lpml, s->dt, s->tab, 0, s->df, a->P, s->L, s->Vxx_left, s->Vxx_right, 1,
                    a->Vxx_left, a->Vxx_right, a->df, gL);
    FN(orc_deriv_t)(a->df, a->Vx, nx, 1, nz, 0, 0, s->order, s->dx);
    FN(orc_deriv)(s->Vz, s->df, nx, 1, nz, 2, 0, s->order, s->dz);
    FN(adj_upd_2d_)(2, nx, nz, s->lpml, s->dt, s->tab, 0, s->df, a->P, s->L, s->Vzz_top, s->Vzz_bottom, 1,
                    a->Vzz_top, a->Vzz_bottom, a->df, gL);
    FN(orc_deriv_t)(a->df, a->Vz, nx, 1, nz, 2, 0, s->order, s->dz);
}

/* Reverse of the velocity half step.  s: forward state with the pressure
 * before the step and the P strips after the velocity half step; a: adjoint
 * state (Vx, Vz = adjoints of the velocities after the half step, unchanged;
 * P and the P strips carried back).  gRx, gRz += dJ/dRx, dJ/dRz. */
void FN(orc_a2d_adj_velocity)(FN(orc_a2d_t) *s, FN(orc_a2d_t) *a, REAL *gRx, REAL *gRz)
{
    int nx = s->nx, nz = s->nz;
    FN(orc_deriv)(s->P, s->df, nx, 1, nz, 0, 1, s->order, s->dx);
    FN(adj_upd_2d_)(0, nx, nz, s->lpml, s->dt, s->tab, 1, s->df, a->Vx, s->Rx, s->P_left, s->P_right, 0,
                    a->P_left, a->P_right, a->df, gRx);
    FN(orc_deriv_t)(a->df, a->P, nx, 1, nz, 0, 1, s->order, s->dx);
    FN(orc_deriv)(s->P, s->df, nx, 1, nz, 2, 1, s->order, s->dz);
    FN(adj_upd_2d_)(2, nx, nz, s->lpml, s->dt, s->tab, 1, s->df, a->Vz, s->Rz, s->P_top, s->P_bottom, 0,
                    a->P_top, a->P_bottom, a->df, gRz);
    FN(orc_deriv_t)(a->df, a->P, nx, 1, nz, 2, 1, s->order, s->dz);
}

/*
 * Bilinear source injection (waves.cpp:407-497, 2085-2171).
 * F += dt*Mod*wav[it+nt*i]*xtri*ytri over the 2x2 corner cells of each
 * in-bounds trace; Mod==NULL gives the elastic explosive form
 * F -= dt*wav*xtri*ytri (applied by the caller to Sxx and Szz).
 * map: int (x,z) pairs, shift: REAL (x,z) pairs, both unpadded grid indices.
 */
void FN(orc_inject2d)(REAL *F, const REAL *Mod, int nx, int lpml, REAL dt,
                      const int *map, const REAL *shift, const REAL *wav,
                      int ntrace, int nt, int it)
{
    for (int i = 0; i < ntrace; i++) {
        if (map[2 * i] >= 0 && map[2 * i + 1] >= 0) {
            REAL xtri[2], ytri[2];
            xtri[0] = 1 - shift[2 * i];
            xtri[1] = shift[2 * i];
            ytri[0] = 1 - shift[2 * i + 1];
            ytri[1] = shift[2 * i + 1];
            for (int i1 = 0; i1 < 2; i1++)
                for (int i2 = 0; i2 < 2; i2++) {
                    long p = (long)(lpml + map[2 * i + 1] + i1) * nx + (lpml + map[2 * i] + i2);
                    REAL w = wav[(long)i * nt + it];
                    if (Mod) F[p] += dt * Mod[p] * w * xtri[i2] * ytri[i1];
                    else F[p] -= dt * w * xtri[i2] * ytri[i1];
                }
        }
    }
}

/* Bilinear receiver gather data[it+nt*i] += xtri*ytri*F (waves.cpp:500-595) */
void FN(orc_record2d)(const REAL *F, int nx, int lpml, const int *map,
                      const REAL *shift, REAL *data, int ntrace, int nt, int it)
{
    for (int i = 0; i < ntrace; i++) {
        if (map[2 * i] >= 0 && map[2 * i + 1] >= 0) {
            REAL xtri[2], ytri[2];
            xtri[0] = 1 - shift[2 * i];
            xtri[1] = shift[2 * i];
            ytri[0] = 1 - shift[2 * i + 1];
            ytri[1] = shift[2 * i + 1];
            for (int i1 = 0; i1 < 2; i1++)
                for (int i2 = 0; i2 < 2; i2++) {
                    long p = (long)(lpml + map[2 * i + 1] + i1) * nx + (lpml + map[2 * i] + i2);
                    data[(long)i * nt + it] += xtri[i2] * ytri[i1] * F[p];
                }
        }
    }
}

/* Elastic pressure gather 0.5*(Sxx+Szz)/factor in double temporaries
 * (waves.cpp:2213-2244).  factor[i] is 1 or the reciprocity modulus. */
void FN(orc_record2d_pressure_el)(const REAL *Sxx, const REAL *Szz, int nx, int lpml,
                                  const int *map, const REAL *shift, const REAL *factor,
                                  REAL *data, int ntrace, int nt, int it)
{
    for (int i = 0; i < ntrace; i++) {
        if (map[2 * i] >= 0 && map[2 * i + 1] >= 0) {
            REAL xtri[2], ytri[2];
            REAL fac = factor[i];
            xtri[0] = 1 - shift[2 * i];
            xtri[1] = shift[2 * i];
            ytri[0] = 1 - shift[2 * i + 1];
            ytri[1] = shift[2 * i + 1];
            for (int i1 = 0; i1 < 2; i1++)
                for (int i2 = 0; i2 < 2; i2++) {
                    long p = (long)(lpml + map[2 * i + 1] + i1) * nx + (lpml + map[2 * i] + i2);
                    long q = (long)i * nt + it;
                    data[q] = (REAL)((double)data[q] + (1. / 2.) * (double)xtri[i2] * (double)ytri[i1] * (double)Sxx[p] / (double)fac);
                    data[q] = (REAL)((double)data[q] + (1. / 2.) * (double)xtri[i2] * (double)ytri[i1] * (double)Szz[p] / (double)fac);
                }
        }
    }
}

/* ---------------------------------------------------------------------- */
/* WavesElastic2D (waves.cpp:1416-2080)                                    */
/* ---------------------------------------------------------------------- */

typedef struct {
    int nx, nz, lpml, order, fs;
    REAL dx, dz, dt;
    const REAL *tab;
    REAL *Sxx, *Szz, *Sxz, *Vx, *Vz;
    /* x-side strips (lpml*nz): Sxx_*, Sxzx_*, Vxx_*, Vzx_*
     * z-side strips (nx*lpml): Szz_*, Sxzz_*, Vzz_*, Vxz_* */
    REAL *Sxx_left, *Sxx_right, *Sxzx_left, *Sxzx_right;
    REAL *Szz_top, *Szz_bottom, *Sxzz_top, *Sxzz_bottom;
    REAL *Vxx_left, *Vxx_right, *Vzx_left, *Vzx_right;
    REAL *Vzz_top, *Vzz_bottom, *Vxz_top, *Vxz_bottom;
    const REAL *L, *L2M, *M, *Rx, *Rz;
    REAL *df, *wrk; /* scratch nx*nz each */
} FN(orc_e2d_t);

void FN(orc_e2d_velocity)(FN(orc_e2d_t) *s)
{
    int nx = s->nx, nz = s->nz, L = s->lpml;
    FN(orc_deriv)(s->Sxx, s->df, nx, 1, nz, 0, 1, s->order, s->dx);
    FN(upd_x2d_)(nx, nz, L, s->dt, s->tab, 1, s->df, s->Vx, s->Rx, 0, 0, s->Sxx_left, s->Sxx_right);
    FN(orc_deriv)(s->Sxz, s->df, nx, 1, nz, 2, 0, s->order, s->dz);
    FN(upd_z2d_)(nx, nz, L, s->dt, s->tab, 0, s->df, s->Vx, s->Rx, 0, 0, s->Sxzz_top, s->Sxzz_bottom);
    FN(orc_deriv)(s->Sxz, s->df, nx, 1, nz, 0, 0, s->order, s->dx);
    FN(upd_x2d_)(nx, nz, L, s->dt, s->tab, 0, s->df, s->Vz, s->Rz, 0, 0, s->Sxzx_left, s->Sxzx_right);
    FN(orc_deriv)(s->Szz, s->df, nx, 1, nz, 2, 1, s->order, s->dz);
    FN(upd_z2d_)(nx, nz, L, s->dt, s->tab, 1, s->df, s->Vz, s->Rz, 0, 0, s->Szz_top, s->Szz_bottom);
}

void FN(orc_e2d_stress)(FN(orc_e2d_t) *s)
{
    int nx = s->nx, nz = s->nz, L = s->lpml;
    FN(orc_deriv)(s->Vx, s->df, nx, 1, nz, 0, 0, s->order, s->dx);
    FN(upd_x2d_)(nx, nz, L, s->dt, s->tab, 0, s->df, s->Sxx, s->L2M, s->Szz, s->L, s->Vxx_left, s->Vxx_right);
    FN(orc_deriv)(s->Vz, s->df, nx, 1, nz, 2, 0, s->order, s->dz);
    FN(upd_z2d_)(nx, nz, L, s->dt, s->tab, 0, s->df, s->Sxx, s->L, s->Szz, s->L2M, s->Vzz_top, s->Vzz_bottom);
    if (s->fs)
        for (int ix = 0; ix < nx; ix++) s->Szz[(long)L * nx + ix] = 0.0;
    FN(orc_deriv)(s->Vz, s->df, nx, 1, nz, 0, 1, s->order, s->dx);
    FN(upd_x2d_)(nx, nz, L, s->dt, s->tab, 1, s->df, s->Sxz, s->M, 0, 0, s->Vzx_left, s->Vzx_right);
    FN(orc_deriv)(s->Vx, s->df, nx, 1, nz, 2, 1, s->order, s->dz);
    FN(upd_z2d_)(nx, nz, L, s->dt, s->tab, 1, s->df, s->Sxz, s->M, 0, 0, s->Vxz_top, s->Vxz_bottom);
}

/* Adjoint-state velocity step (waves.cpp:1738-1904): derivatives of the
 * moduli-weighted stresses. */
void FN(orc_e2d_velocity_bw)(FN(orc_e2d_t) *s)
{
    int nx = s->nx, nz = s->nz, L = s->lpml;
    long n = (long)nx * nz;
    for (long p = 0; p < n; p++) s->wrk[p] = s->L2M[p] * s->Sxx[p] + s->L[p] * s->Szz[p];
    FN(orc_deriv)(s->wrk, s->df, nx, 1, nz, 0, 1, s->order, s->dx);
    FN(upd_x2d_)(nx, nz, L, s->dt, s->tab, 1, s->df, s->Vx, s->Rx, 0, 0, s->Sxx_left, s->Sxx_right);
    for (long p = 0; p < n; p++) s->wrk[p] = s->M[p] * s->Sxz[p];
    FN(orc_deriv)(s->wrk, s->df, nx, 1, nz, 2, 0, s->order, s->dz);
    FN(upd_z2d_)(nx, nz, L, s->dt, s->tab, 0, s->df, s->Vx, s->Rx, 0, 0, s->Sxzz_top, s->Sxzz_bottom);
    FN(orc_deriv)(s->wrk, s->df, nx, 1, nz, 0, 0, s->order, s->dx);
    FN(upd_x2d_)(nx, nz, L, s->dt, s->tab, 0, s->df, s->Vz, s->Rz, 0, 0, s->Sxzx_left, s->Sxzx_right);
    for (long p = 0; p < n; p++) s->wrk[p] = s->L[p] * s->Sxx[p] + s->L2M[p] * s->Szz[p];
    FN(orc_deriv)(s->wrk, s->df, nx, 1, nz, 2, 1, s->order, s->dz);
    FN(upd_z2d_)(nx, nz, L, s->dt, s->tab, 1, s->df, s->Vz, s->Rz, 0, 0, s->Szz_top, s->Szz_bottom);
}

/* Adjoint-state stress step (waves.cpp:1907-2080): bare dt*dV, no moduli. */
void FN(orc_e2d_stress_bw)(FN(orc_e2d_t) *s)
{
    int nx = s->nx, nz = s->nz, L = s->lpml;
    FN(orc_deriv)(s->Vx, s->df, nx, 1, nz, 0, 0, s->order, s->dx);
    FN(upd_x2d_)(nx, nz, L, s->dt, s->tab, 0, s->df, s->Sxx, 0, 0, 0, s->Vxx_left, s->Vxx_right);
    FN(orc_deriv)(s->Vz, s->df, nx, 1, nz, 2, 0, s->order, s->dz);
    FN(upd_z2d_)(nx, nz, L, s->dt, s->tab, 0, s->df, s->Szz, 0, 0, 0, s->Vzz_top, s->Vzz_bottom);
    if (s->fs)
        for (int ix = 0; ix < nx; ix++) s->Szz[(long)L * nx + ix] = 0.0;
    FN(orc_deriv)(s->Vz, s->df, nx, 1, nz, 0, 1, s->order, s->dx);
    FN(upd_x2d_)(nx, nz, L, s->dt, s->tab, 1, s->df, s->Sxz, 0, 0, 0, s->Vzx_left, s->Vzx_right);
    FN(orc_deriv)(s->Vx, s->df, nx, 1, nz, 2, 1, s->order, s->dz);
    FN(upd_z2d_)(nx, nz, L, s->dt, s->tab, 1, s->df, s->Sxz, 0, 0, 0, s->Vxz_top, s->Vxz_bottom);
}

/* ---------------------------------------------------------------------- */
/* Exact discrete adjoint of the elastic 2D step (no reference counterpart: */
/* FwiElastic2D back-propagates with the adjoint-state steps above, which   */
/* are not the transpose of the discrete forward step).  Every linear map   */
/* of orc_e2d_velocity / orc_e2d_stress is transposed as it stands.         */
/* ---------------------------------------------------------------------- */

/*
 * Reverse of one derivative term of upd_x2d_ (axis 0) / upd_z2d_ (axis 2)
 * with up to two targets sharing the strip: F += dt*M*d, F2 += dt*M2*d
 * (outside the strips), minus dt*M*(psi' + C*d) (and the same for F2) inside,
 * psi' = B*psi + A*d.  lF, lF2 are the adjoints of F', F2' (lF2 NULL: one
 * target).  With S = M*lF + M2*lF2 this writes w = adjoint of d, replaces
 * lpsi by the adjoint of the old psi and adds lF * dF'/dM to g and
 * lF2 * dF2'/dM2 to g2.  psi holds the forward strip before (psi_old=1) or
 * after (psi_old=0) the update.
 */
static void FN(adj_upd2_2d_)(int axis, int nx, int nz, int L, REAL dt, const REAL *tab, int stag,
                             const REAL *df, const REAL *lF, const REAL *M, const REAL *lF2, const REAL *M2,
                             const REAL *psiLo, const REAL *psiHi, int psi_old, REAL *lpsiLo, REAL *lpsiHi,
                             REAL *w, REAL *g, REAL *g2)
{
    long n = (long)nx * nz;
    for (long p = 0; p < n; p++) {
        REAL S = M[p] * lF[p] + (lF2 ? M2[p] * lF2[p] : 0.0);
        w[p] = dt * S;
        g[p] += lF[p] * dt * df[p];
        if (lF2) g2[p] += lF2[p] * dt * df[p];
    }
    int n1 = axis == 0 ? nz : L, n2 = axis == 0 ? L : nx;
    for (int side = 0; side < 2; side++) {
        const REAL *At = side ? TAB_A_R(tab, L, stag) : TAB_A_L(tab, L, stag);
        const REAL *Bt = side ? TAB_B_R(tab, L, stag) : TAB_B_L(tab, L, stag);
        const REAL *Ct = side ? TAB_C_R(tab, L, stag) : TAB_C_L(tab, L, stag);
        const REAL *psi = side ? psiHi : psiLo;
        REAL *lpsi = side ? lpsiHi : lpsiLo;
        for (int i1 = 0; i1 < n1; i1++)
            for (int i2 = 0; i2 < n2; i2++) {
                long q = (long)i1 * n2 + i2, p;
                int t;
                if (axis == 0) {
                    t = i2;
                    p = (long)i1 * nx + i2 + (side ? nx - L : 0);
                } else {
                    t = i1;
                    p = (long)(i1 + (side ? nz - L : 0)) * nx + i2;
                }
                REAL A = At[t], B = Bt[t], C = Ct[t];
                REAL pn = psi_old ? B * psi[q] + A * df[p] : psi[q];
                REAL S = M[p] * lF[p] + (lF2 ? M2[p] * lF2[p] : 0.0);
                REAL lq = lpsi[q], dtS = dt * S, e = dt * (pn + C * df[p]);
                w[p] = dtS * (1 - A - C) + A * lq;
                lpsi[q] = B * lq - dtS * B;
                g[p] -= lF[p] * e;
                if (lF2) g2[p] -= lF2[p] * e;
            }
    }
}

/* Reverse of the stress half step.  s: forward state holding V of this
 * step (after the velocity half step) and the V strips before the stress
 * half step; a: adjoint state (stresses = adjoints after the step; the
 * velocities and V strips carried back).  The free surface zeroes Szz on
 * row lpml after the normal-stress update, so its adjoint is dropped there
 * first.  gL, gL2M, gM += the gradients of the moduli. */
void FN(orc_e2d_adj_stress)(FN(orc_e2d_t) *s, FN(orc_e2d_t) *a, REAL *gL, REAL *gL2M, REAL *gM)
{
    int nx = s->nx, nz = s->nz, L = s->lpml;
    if (s->fs)
        for (int ix = 0; ix < nx; ix++) a->Szz[(long)L * nx + ix] = 0.0;
    FN(orc_deriv)(s->Vx, s->df, nx, 1, nz, 0, 0, s->order, s->dx);
    FN(adj_upd2_2d_)(0, nx, nz, L, s->dt, s->tab, 0, s->df, a->Sxx, s->L2M, a->Szz, s->L, s->Vxx_left,
                     s->Vxx_right, 1, a->Vxx_left, a->Vxx_right, a->df, gL2M, gL);
    FN(orc_deriv_t)(a->df, a->Vx, nx, 1, nz, 0, 0, s->order, s->dx);
    FN(orc_deriv)(s->Vz, s->df, nx, 1, nz, 2, 0, s->order, s->dz);
    FN(adj_upd2_2d_)(2, nx, nz, L, s->dt, s->tab, 0, s->df, a->Sxx, s->L, a->Szz, s->L2M, s->Vzz_top,
                     s->Vzz_bottom, 1, a->Vzz_top, a->Vzz_bottom, a->df, gL, gL2M);
    FN(orc_deriv_t)(a->df, a->Vz, nx, 1, nz, 2, 0, s->order, s->dz);
    FN(orc_deriv)(s->Vz, s->df, nx, 1, nz, 0, 1, s->order, s->dx);
    FN(adj_upd2_2d_)(0, nx, nz, L, s->dt, s->tab, 1, s->df, a->Sxz, s->M, 0, 0, s->Vzx_left, s->Vzx_right, 1,
                     a->Vzx_left, a->Vzx_right, a->df, gM, 0);
    FN(orc_deriv_t)(a->df, a->Vz, nx, 1, nz, 0, 1, s->order, s->dx);
    FN(orc_deriv)(s->Vx, s->df, nx, 1, nz, 2, 1, s->order, s->dz);
    FN(adj_upd2_2d_)(2, nx, nz, L, s->dt, s->tab, 1, s->df, a->Sxz, s->M, 0, 0, s->Vxz_top, s->Vxz_bottom, 1,
                     a->Vxz_top, a->Vxz_bottom, a->df, gM, 0);
    FN(orc_deriv_t)(a->df, a->Vx, nx, 1, nz, 2, 1, s->order, s->dz);
}

/* Reverse of the velocity half step.  s: forward state with the stresses
 * before the step and the stress strips after the velocity half step; a:
 * adjoint state (velocities = adjoints after the half step, unchanged;
 * stresses and stress strips carried back).  gRx, gRz += dJ/dRx, dJ/dRz. */
void FN(orc_e2d_adj_velocity)(FN(orc_e2d_t) *s, FN(orc_e2d_t) *a, REAL *gRx, REAL *gRz)
{
    int nx = s->nx, nz = s->nz, L = s->lpml;
    FN(orc_deriv)(s->Sxx, s->df, nx, 1, nz, 0, 1, s->order, s->dx);
    FN(adj_upd2_2d_)(0, nx, nz, L, s->dt, s->tab, 1, s->df, a->Vx, s->Rx, 0, 0, s->Sxx_left, s->Sxx_right, 0,
                     a->Sxx_left, a->Sxx_right, a->df, gRx, 0);
    FN(orc_deriv_t)(a->df, a->Sxx, nx, 1, nz, 0, 1, s->order, s->dx);
    FN(orc_deriv)(s->Sxz, s->df, nx, 1, nz, 2, 0, s->order, s->dz);
    FN(adj_upd2_2d_)(2, nx, nz, L, s->dt, s->tab, 0, s->df, a->Vx, s->Rx, 0, 0, s->Sxzz_top, s->Sxzz_bottom, 0,
                     a->Sxzz_top, a->Sxzz_bottom, a->df, gRx, 0);
    FN(orc_deriv_t)(a->df, a->Sxz, nx, 1, nz, 2, 0, s->order, s->dz);
    FN(orc_deriv)(s->Sxz, s->df, nx, 1, nz, 0, 0, s->order, s->dx);
    FN(adj_upd2_2d_)(0, nx, nz, L, s->dt, s->tab, 0, s->df, a->Vz, s->Rz, 0, 0, s->Sxzx_left, s->Sxzx_right, 0,
                     a->Sxzx_left, a->Sxzx_right, a->df, gRz, 0);
    FN(orc_deriv_t)(a->df, a->Sxz, nx, 1, nz, 0, 0, s->order, s->dx);
    FN(orc_deriv)(s->Szz, s->df, nx, 1, nz, 2, 1, s->order, s->dz);
    FN(adj_upd2_2d_)(2, nx, nz, L, s->dt, s->tab, 1, s->df, a->Vz, s->Rz, 0, 0, s->Szz_top, s->Szz_bottom, 0,
                     a->Szz_top, a->Szz_bottom, a->df, gRz, 0);
    FN(orc_deriv_t)(a->df, a->Szz, nx, 1, nz, 2, 1, s->order, s->dz);
}

/* ---------------------------------------------------------------------- */
/* WavesAcoustic3D (waves.cpp:710-1275)                                    */
/* ---------------------------------------------------------------------- */

typedef struct {
    int nx, ny, nz, lpml, order;
    REAL dx, dy, dz, dt;
    const REAL *tab;
    REAL *P, *Vx, *Vy, *Vz;
    /* P_left,P_right (lpml*ny*nz) P_front,P_back (nx*lpml*nz)
     * P_top,P_bottom (nx*ny*lpml); same for Vxx_*, Vyy_*, Vzz_* */
    REAL *P_left, *P_right, *P_front, *P_back, *P_top, *P_bottom;
    REAL *Vxx_left, *Vxx_right, *Vyy_front, *Vyy_back, *Vzz_top, *Vzz_bottom;
    const REAL *L, *Rx, *Ry, *Rz;
    REAL *df;
} FN(orc_a3d_t);

/* F += dt*M*df over the grid, then the two strips of `axis`. */
static void FN(upd3d_)(FN(orc_a3d_t) *s, int axis, int stag, REAL *F, const REAL *M,
                       REAL *psiLo, REAL *psiHi)
{
    int nx = s->nx, ny = s->ny, nz = s->nz, L = s->lpml;
    REAL dt = s->dt;
    const REAL *df = s->df, *tab = s->tab;
    long nxy = (long)nx * ny, n = nxy * nz;
    for (long p = 0; p < n; p++) F[p] += dt * M[p] * df[p];
    const REAL *Al = TAB_A_L(tab, L, stag), *Bl = TAB_B_L(tab, L, stag), *Cl = TAB_C_L(tab, L, stag);
    const REAL *Ar = TAB_A_R(tab, L, stag), *Br = TAB_B_R(tab, L, stag), *Cr = TAB_C_R(tab, L, stag);
    int ex = axis == 0 ? L : nx, ey = axis == 1 ? L : ny, ez = axis == 2 ? L : nz;
    for (int iz = 0; iz < ez; iz++)
        for (int iy = 0; iy < ey; iy++)
            for (int ix = 0; ix < ex; ix++) {
                long q, p, ph;
                int t;
                if (axis == 0) {
                    q = (long)iz * L * ny + (long)iy * L + ix;
                    p = iz * nxy + (long)iy * nx + ix;
                    ph = p + nx - L;
                    t = ix;
                } else if (axis == 1) {
                    q = (long)iz * nx * L + (long)iy * nx + ix;
                    p = iz * nxy + (long)iy * nx + ix;
                    ph = p + (long)(ny - L) * nx;
                    t = iy;
                } else {
                    q = iz * nxy + (long)iy * nx + ix;
                    p = q;
                    ph = p + (long)(nz - L) * nxy;
                    t = iz;
                }
                PSI_UPD(psiLo[q], Bl[t], Al[t], Cl[t], df[p], F[p], M[p], dt);
                PSI_UPD(psiHi[q], Br[t], Ar[t], Cr[t], df[ph], F[ph], M[ph], dt);
            }
}

void FN(orc_a3d_velocity)(FN(orc_a3d_t) *s)
{
    FN(orc_deriv)(s->P, s->df, s->nx, s->ny, s->nz, 0, 1, s->order, s->dx);
    FN(upd3d_)(s, 0, 1, s->Vx, s->Rx, s->P_left, s->P_right);
    FN(orc_deriv)(s->P, s->df, s->nx, s->ny, s->nz, 1, 1, s->order, s->dy);
    FN(upd3d_)(s, 1, 1, s->Vy, s->Ry, s->P_front, s->P_back);
    FN(orc_deriv)(s->P, s->df, s->nx, s->ny, s->nz, 2, 1, s->order, s->dz);
    FN(upd3d_)(s, 2, 1, s->Vz, s->Rz, s->P_top, s->P_bottom);
}

void FN(orc_a3d_stress)(FN(orc_a3d_t) *s)
{
    FN(orc_deriv)(s->Vx, s->df, s->nx, s->ny, s->nz, 0, 0, s->order, s->dx);
    FN(upd3d_)(s, 0, 0, s->P, s->L, s->Vxx_left, s->Vxx_right);
    FN(orc_deriv)(s->Vy, s->df, s->nx, s->ny, s->nz, 1, 0, s->order, s->dy);
    FN(upd3d_)(s, 1, 0, s->P, s->L, s->Vyy_front, s->Vyy_back);
    FN(orc_deriv)(s->Vz, s->df, s->nx, s->ny, s->nz, 2, 0, s->order, s->dz);
    FN(upd3d_)(s, 2, 0, s->P, s->L, s->Vzz_top, s->Vzz_bottom);
}

/* Exact discrete adjoint of the acoustic 3D step (no reference counterpart:
 * FwiAcoustic3D back-propagates with the forward operator).  Reverse of one
 * derivative term of upd3d_: F' = F + dt*M*d outside the strips, inside
 * psi' = B*psi + A*d and F' -= dt*M*(psi' + C*d).  lF = adjoint of F',
 * lpsi = adjoint of psi' (replaced by the adjoint of the old psi), w = the
 * adjoint of d, g += lF dF'/dM; psi holds the forward strip before
 * (psi_old=1) or after (psi_old=0) the update; d in s->df. */
static void FN(adj_upd3d_)(FN(orc_a3d_t) *s, int axis, int stag, const REAL *lF, const REAL *M, const REAL *psiLo,
                           const REAL *psiHi, int psi_old, REAL *lpsiLo, REAL *lpsiHi, REAL *w, REAL *g)
{
    int nx = s->nx, ny = s->ny, nz = s->nz, L = s->lpml;
    REAL dt = s->dt;
    const REAL *df = s->df, *tab = s->tab;
    long nxy = (long)nx * ny, n = nxy * nz;
    for (long p = 0; p < n; p++) {
        w[p] = dt * M[p] * lF[p];
        g[p] += lF[p] * dt * df[p];
    }
    int ex = axis == 0 ? L : nx, ey = axis == 1 ? L : ny, ez = axis == 2 ? L : nz;
    for (int side = 0; side < 2; side++) {
        const REAL *At = side ? TAB_A_R(tab, L, stag) : TAB_A_L(tab, L, stag);
        const REAL *Bt = side ? TAB_B_R(tab, L, stag) : TAB_B_L(tab, L, stag);
        const REAL *Ct = side ? TAB_C_R(tab, L, stag) : TAB_C_L(tab, L, stag);
        const REAL *psi = side ? psiHi : psiLo;
        REAL *lpsi = side ? lpsiHi : lpsiLo;
        for (int iz = 0; iz < ez; iz++)
            for (int iy = 0; iy < ey; iy++)
                for (int ix = 0; ix < ex; ix++) {
                    long q, p;
                    int t;
                    if (axis == 0) {
                        q = (long)iz * L * ny + (long)iy * L + ix;
                        p = iz * nxy + (long)iy * nx + ix + (side ? nx - L : 0);
                        t = ix;
                    } else if (axis == 1) {
                        q = (long)iz * nx * L + (long)iy * nx + ix;
                        p = iz * nxy + (long)(iy + (side ? ny - L : 0)) * nx + ix;
                        t = iy;
                    } else {
                        q = iz * nxy + (long)iy * nx + ix;
                        p = (iz + (side ? nz - L : 0)) * nxy + (long)iy * nx + ix;
                        t = iz;
                    }
                    REAL A = At[t], B = Bt[t], C = Ct[t];
                    REAL pn = psi_old ? B * psi[q] + A * df[p] : psi[q];
                    REAL lf = lF[p], lq = lpsi[q], dtM = dt * M[p];
                    w[p] = dtM * (1 - A - C) * lf + A * lq;
                    lpsi[q] = B * lq - dtM * B * lf;
                    g[p] -= lf * dt * (pn + C * df[p]);
                }
    }
}

/* Reverse of the stress half step: s holds V of this step and the Vxx /
 * Vyy / Vzz strips before the stress half step; a: adjoint state (P =
 * adjoint of the pressure after the step, unchanged; velocities and V
 * strips carried back).  gL += dJ/dL. */
void FN(orc_a3d_adj_stress)(FN(orc_a3d_t) *s, FN(orc_a3d_t) *a, REAL *gL)
{
    int nx = s->nx, ny = s->ny, nz = s->nz;
    REAL *V[3] = {s->Vx, s->Vy, s->Vz}, *lV[3] = {a->Vx, a->Vy, a->Vz};
    REAL *pl[3] = {s->Vxx_left, s->Vyy_front, s->Vzz_top}, *ph[3] = {s->Vxx_right, s->Vyy_back, s->Vzz_bottom};
    REAL *ll[3] = {a->Vxx_left, a->Vyy_front, a->Vzz_top}, *lh[3] = {a->Vxx_right, a->Vyy_back, a->Vzz_bottom};
    REAL d[3] = {s->dx, s->dy, s->dz};
    for (int ax = 0; ax < 3; ax++) {
        FN(orc_deriv)(V[ax], s->df, nx, ny, nz, ax, 0, s->order, d[ax]);
        FN(adj_upd3d_)(s, ax, 0, a->P, s->L, pl[ax], ph[ax], 1, ll[ax], lh[ax], a->df, gL);
        FN(orc_deriv_t)(a->df, lV[ax], nx, ny, nz, ax, 0, s->order, d[ax]);
    }
}

/* Reverse of the velocity half step: s holds P before the step and the P
 * strips after the velocity half step; a: adjoint state (velocities =
 * adjoints after the half step, unchanged; P and the P strips carried
 * back).  gRx, gRy, gRz += dJ/dRx, dJ/dRy, dJ/dRz. */
void FN(orc_a3d_adj_velocity)(FN(orc_a3d_t) *s, FN(orc_a3d_t) *a, REAL *gRx, REAL *gRy, REAL *gRz)
{
    int nx = s->nx, ny = s->ny, nz = s->nz;
    REAL *lV[3] = {a->Vx, a->Vy, a->Vz}, *g[3] = {gRx, gRy, gRz};
    const REAL *R[3] = {s->Rx, s->Ry, s->Rz};
    REAL *pl[3] = {s->P_left, s->P_front, s->P_top}, *ph[3] = {s->P_right, s->P_back, s->P_bottom};
    REAL *ll[3] = {a->P_left, a->P_front, a->P_top}, *lh[3] = {a->P_right, a->P_back, a->P_bottom};
    REAL d[3] = {s->dx, s->dy, s->dz};
    for (int ax = 0; ax < 3; ax++) {
        FN(orc_deriv)(s->P, s->df, nx, ny, nz, ax, 1, s->order, d[ax]);
        FN(adj_upd3d_)(s, ax, 1, lV[ax], R[ax], pl[ax], ph[ax], 0, ll[ax], lh[ax], a->df, g[ax]);
        FN(orc_deriv_t)(a->df, a->P, nx, ny, nz, ax, 1, s->order, d[ax]);
    }
}

/* Trilinear injection / gather (waves.cpp:998-1275); map/shift (x,y,z). */
void FN(orc_inject3d)(REAL *F, const REAL *Mod, int nx, int ny, int lpml, REAL dt,
                      const int *map, const REAL *shift, const REAL *wav,
                      int ntrace, int nt, int it)
{
    long nxy = (long)nx * ny;
    for (int i = 0; i < ntrace; i++) {
        const int *m = map + 3 * i;
        const REAL *sh = shift + 3 * i;
        if (m[0] >= 0 && m[1] >= 0 && m[2] >= 0) {
            REAL xtri[2] = {1 - sh[0], sh[0]}, ytri[2] = {1 - sh[1], sh[1]}, ztri[2] = {1 - sh[2], sh[2]};
            for (int i1 = 0; i1 < 2; i1++)
                for (int i2 = 0; i2 < 2; i2++)
                    for (int i3 = 0; i3 < 2; i3++) {
                        long p = (long)(lpml + m[2] + i1) * nxy + (long)(lpml + m[1] + i2) * nx + (lpml + m[0] + i3);
                        F[p] += dt * Mod[p] * wav[(long)i * nt + it] * xtri[i3] * ytri[i2] * ztri[i1];
                    }
        }
    }
}

void FN(orc_record3d)(const REAL *F, int nx, int ny, int lpml, const int *map,
                      const REAL *shift, REAL *data, int ntrace, int nt, int it)
{
    long nxy = (long)nx * ny;
    for (int i = 0; i < ntrace; i++) {
        const int *m = map + 3 * i;
        const REAL *sh = shift + 3 * i;
        if (m[0] >= 0 && m[1] >= 0 && m[2] >= 0) {
            REAL xtri[2] = {1 - sh[0], sh[0]}, ytri[2] = {1 - sh[1], sh[1]}, ztri[2] = {1 - sh[2], sh[2]};
            for (int i1 = 0; i1 < 2; i1++)
                for (int i2 = 0; i2 < 2; i2++)
                    for (int i3 = 0; i3 < 2; i3++) {
                        long p = (long)(lpml + m[2] + i1) * nxy + (long)(lpml + m[1] + i2) * nx + (lpml + m[0] + i3);
                        data[(long)i * nt + it] += xtri[i3] * ytri[i2] * ztri[i1] * F[p];
                    }
        }
    }
}

/* ---------------------------------------------------------------------- */
/* WavesElastic3D (waves.cpp:2310-3258, 4104-4403)                          */
/* ---------------------------------------------------------------------- */

/* Strip slots (PmlElastic3D, pml.h:197-234): x faces (lpml*ny*nz each)
 * 0..11 Sxx, Sxzx, Sxyx, Vxx, Vzx, Vyx (left, right); y faces (nx*lpml*nz)
 * 12..23 Syy, Sxyy, Syzy, Vyy, Vxy, Vzy (front, back); z faces (nx*ny*lpml)
 * 24..35 Szz, Sxzz, Syzz, Vzz, Vxz, Vyz (top, bottom). */
typedef struct {
    int nx, ny, nz, lpml, order, fs;
    REAL dx, dy, dz, dt;
    const REAL *tab;
    REAL *Vx, *Vy, *Vz, *Sxx, *Syy, *Szz, *Syz, *Sxz, *Sxy;
    REAL *psi[36];
    const REAL *L, *L2M, *Mxz, *Myz, *Mxy, *Rx, *Ry, *Rz;
    REAL *df;
    REAL *wrk; /* backward velocity step: the moduli-weighted stresses */
} FN(orc_e3d_t);

/* F_k += dt*M_k*df over the grid (k = 1..3, F2/F3 optional; M1 NULL: the
 * backward stress step's bare dt*df), then the two strips of `axis`:
 * psi = B*psi + A*df, F_k -= dt*M_k*(psi + C*df). */
static void FN(upd3e_)(FN(orc_e3d_t) *s, int axis, int stag, REAL *F1, const REAL *M1, REAL *F2, const REAL *M2,
                       REAL *F3, const REAL *M3, int slot)
{
    int nx = s->nx, ny = s->ny, nz = s->nz, L = s->lpml;
    REAL dt = s->dt;
    const REAL *df = s->df, *tab = s->tab;
    REAL *psiLo = s->psi[slot], *psiHi = s->psi[slot + 1];
    long nxy = (long)nx * ny, n = nxy * nz;
    for (long p = 0; p < n; p++) {
        F1[p] += M1 ? dt * M1[p] * df[p] : dt * df[p];
        if (F2) F2[p] += dt * M2[p] * df[p];
        if (F3) F3[p] += dt * M3[p] * df[p];
    }
    const REAL *Al = TAB_A_L(tab, L, stag), *Bl = TAB_B_L(tab, L, stag), *Cl = TAB_C_L(tab, L, stag);
    const REAL *Ar = TAB_A_R(tab, L, stag), *Br = TAB_B_R(tab, L, stag), *Cr = TAB_C_R(tab, L, stag);
    int ex = axis == 0 ? L : nx, ey = axis == 1 ? L : ny, ez = axis == 2 ? L : nz;
    for (int iz = 0; iz < ez; iz++)
        for (int iy = 0; iy < ey; iy++)
            for (int ix = 0; ix < ex; ix++) {
                long q, p, ph;
                int t;
                if (axis == 0) {
                    q = (long)iz * L * ny + (long)iy * L + ix;
                    p = iz * nxy + (long)iy * nx + ix;
                    ph = p + nx - L;
                    t = ix;
                } else if (axis == 1) {
                    q = (long)iz * nx * L + (long)iy * nx + ix;
                    p = iz * nxy + (long)iy * nx + ix;
                    ph = p + (long)(ny - L) * nx;
                    t = iy;
                } else {
                    q = iz * nxy + (long)iy * nx + ix;
                    p = q;
                    ph = p + (long)(nz - L) * nxy;
                    t = iz;
                }
                for (int side = 0; side < 2; side++) {
                    long c = side ? ph : p;
                    REAL *psi = side ? psiHi : psiLo;
                    REAL A = side ? Ar[t] : Al[t], B = side ? Br[t] : Bl[t], C = side ? Cr[t] : Cl[t];
                    psi[q] = B * psi[q] + A * df[c];
                    F1[c] -= M1 ? dt * M1[c] * (psi[q] + C * df[c]) : dt * (psi[q] + C * df[c]);
                    if (F2) F2[c] -= dt * M2[c] * (psi[q] + C * df[c]);
                    if (F3) F3[c] -= dt * M3[c] * (psi[q] + C * df[c]);
                }
            }
}

void FN(orc_e3d_velocity)(FN(orc_e3d_t) *s)
{
    int nx = s->nx, ny = s->ny, nz = s->nz, o = s->order;
    FN(orc_deriv)(s->Sxx, s->df, nx, ny, nz, 0, 1, o, s->dx);
    FN(upd3e_)(s, 0, 1, s->Vx, s->Rx, 0, 0, 0, 0, 0);
    FN(orc_deriv)(s->Sxy, s->df, nx, ny, nz, 1, 0, o, s->dy);
    FN(upd3e_)(s, 1, 0, s->Vx, s->Rx, 0, 0, 0, 0, 14);
    FN(orc_deriv)(s->Sxz, s->df, nx, ny, nz, 2, 0, o, s->dz);
    FN(upd3e_)(s, 2, 0, s->Vx, s->Rx, 0, 0, 0, 0, 26);
    FN(orc_deriv)(s->Sxy, s->df, nx, ny, nz, 0, 0, o, s->dx);
    FN(upd3e_)(s, 0, 0, s->Vy, s->Ry, 0, 0, 0, 0, 4);
    FN(orc_deriv)(s->Syy, s->df, nx, ny, nz, 1, 1, o, s->dy);
    FN(upd3e_)(s, 1, 1, s->Vy, s->Ry, 0, 0, 0, 0, 12);
    FN(orc_deriv)(s->Syz, s->df, nx, ny, nz, 2, 0, o, s->dz);
    FN(upd3e_)(s, 2, 0, s->Vy, s->Ry, 0, 0, 0, 0, 28);
    FN(orc_deriv)(s->Sxz, s->df, nx, ny, nz, 0, 0, o, s->dx);
    FN(upd3e_)(s, 0, 0, s->Vz, s->Rz, 0, 0, 0, 0, 2);
    FN(orc_deriv)(s->Syz, s->df, nx, ny, nz, 1, 0, o, s->dy);
    FN(upd3e_)(s, 1, 0, s->Vz, s->Rz, 0, 0, 0, 0, 16);
    FN(orc_deriv)(s->Szz, s->df, nx, ny, nz, 2, 1, o, s->dz);
    FN(upd3e_)(s, 2, 1, s->Vz, s->Rz, 0, 0, 0, 0, 24);
}

void FN(orc_e3d_stress)(FN(orc_e3d_t) *s)
{
    int nx = s->nx, ny = s->ny, nz = s->nz, o = s->order;
    long nxy = (long)nx * ny;
    FN(orc_deriv)(s->Vx, s->df, nx, ny, nz, 0, 0, o, s->dx);
    FN(upd3e_)(s, 0, 0, s->Sxx, s->L2M, s->Syy, s->L, s->Szz, s->L, 6);
    FN(orc_deriv)(s->Vy, s->df, nx, ny, nz, 1, 0, o, s->dy);
    FN(upd3e_)(s, 1, 0, s->Sxx, s->L, s->Syy, s->L2M, s->Szz, s->L, 18);
    FN(orc_deriv)(s->Vz, s->df, nx, ny, nz, 2, 0, o, s->dz);
    /* the free surface row lpml lies in no z strip: zeroing Szz there after
     * the strips equals the reference's zeroing before them */
    FN(upd3e_)(s, 2, 0, s->Sxx, s->L, s->Syy, s->L, s->Szz, s->L2M, 30);
    if (s->fs)
        for (long p = 0; p < nxy; p++) s->Szz[(long)s->lpml * nxy + p] = 0.0;
    FN(orc_deriv)(s->Vz, s->df, nx, ny, nz, 0, 1, o, s->dx);
    FN(upd3e_)(s, 0, 1, s->Sxz, s->Mxz, 0, 0, 0, 0, 8);
    FN(orc_deriv)(s->Vx, s->df, nx, ny, nz, 2, 1, o, s->dz);
    FN(upd3e_)(s, 2, 1, s->Sxz, s->Mxz, 0, 0, 0, 0, 32);
    FN(orc_deriv)(s->Vy, s->df, nx, ny, nz, 2, 1, o, s->dz);
    FN(upd3e_)(s, 2, 1, s->Syz, s->Myz, 0, 0, 0, 0, 34);
    FN(orc_deriv)(s->Vz, s->df, nx, ny, nz, 1, 1, o, s->dy);
    FN(upd3e_)(s, 1, 1, s->Syz, s->Myz, 0, 0, 0, 0, 22);
    FN(orc_deriv)(s->Vx, s->df, nx, ny, nz, 1, 1, o, s->dy);
    FN(upd3e_)(s, 1, 1, s->Sxy, s->Mxy, 0, 0, 0, 0, 20);
    FN(orc_deriv)(s->Vy, s->df, nx, ny, nz, 0, 1, o, s->dx);
    FN(upd3e_)(s, 0, 1, s->Sxy, s->Mxy, 0, 0, 0, 0, 10);
}

/* wrk = M1*F1 (+ M2*F2 + M3*F3) cell by cell, the reference's wrk arrays of
 * backwardstepVelocity. */
static void FN(wrk3_)(FN(orc_e3d_t) *s, const REAL *M1, const REAL *F1, const REAL *M2, const REAL *F2,
                      const REAL *M3, const REAL *F3)
{
    long n = (long)s->nx * s->ny * s->nz;
    for (long p = 0; p < n; p++)
        s->wrk[p] = F2 ? M1[p] * F1[p] + M2[p] * F2[p] + M3[p] * F3[p] : M1[p] * F1[p];
}

/* WavesElastic3D::backwardstepVelocity (waves.cpp:3260-3713): the forward
 * velocity terms and strips on wrk = L2M Sxx + L Syy + L Szz (x), M_xy Sxy,
 * M_xz Sxz, ... */
void FN(orc_e3d_velocity_bw)(FN(orc_e3d_t) *s)
{
    int nx = s->nx, ny = s->ny, nz = s->nz, o = s->order;
    FN(wrk3_)(s, s->L2M, s->Sxx, s->L, s->Syy, s->L, s->Szz);
    FN(orc_deriv)(s->wrk, s->df, nx, ny, nz, 0, 1, o, s->dx);
    FN(upd3e_)(s, 0, 1, s->Vx, s->Rx, 0, 0, 0, 0, 0);
    FN(wrk3_)(s, s->Mxy, s->Sxy, 0, 0, 0, 0);
    FN(orc_deriv)(s->wrk, s->df, nx, ny, nz, 1, 0, o, s->dy);
    FN(upd3e_)(s, 1, 0, s->Vx, s->Rx, 0, 0, 0, 0, 14);
    FN(wrk3_)(s, s->Mxz, s->Sxz, 0, 0, 0, 0);
    FN(orc_deriv)(s->wrk, s->df, nx, ny, nz, 2, 0, o, s->dz);
    FN(upd3e_)(s, 2, 0, s->Vx, s->Rx, 0, 0, 0, 0, 26);
    FN(wrk3_)(s, s->Mxy, s->Sxy, 0, 0, 0, 0);
    FN(orc_deriv)(s->wrk, s->df, nx, ny, nz, 0, 0, o, s->dx);
    FN(upd3e_)(s, 0, 0, s->Vy, s->Ry, 0, 0, 0, 0, 4);
    FN(wrk3_)(s, s->L, s->Sxx, s->L2M, s->Syy, s->L, s->Szz);
    FN(orc_deriv)(s->wrk, s->df, nx, ny, nz, 1, 1, o, s->dy);
    FN(upd3e_)(s, 1, 1, s->Vy, s->Ry, 0, 0, 0, 0, 12);
    FN(wrk3_)(s, s->Myz, s->Syz, 0, 0, 0, 0);
    FN(orc_deriv)(s->wrk, s->df, nx, ny, nz, 2, 0, o, s->dz);
    FN(upd3e_)(s, 2, 0, s->Vy, s->Ry, 0, 0, 0, 0, 28);
    FN(wrk3_)(s, s->Mxz, s->Sxz, 0, 0, 0, 0);
    FN(orc_deriv)(s->wrk, s->df, nx, ny, nz, 0, 0, o, s->dx);
    FN(upd3e_)(s, 0, 0, s->Vz, s->Rz, 0, 0, 0, 0, 2);
    FN(wrk3_)(s, s->Myz, s->Syz, 0, 0, 0, 0);
    FN(orc_deriv)(s->wrk, s->df, nx, ny, nz, 1, 0, o, s->dy);
    FN(upd3e_)(s, 1, 0, s->Vz, s->Rz, 0, 0, 0, 0, 16);
    FN(wrk3_)(s, s->L, s->Sxx, s->L, s->Syy, s->L2M, s->Szz);
    FN(orc_deriv)(s->wrk, s->df, nx, ny, nz, 2, 1, o, s->dz);
    FN(upd3e_)(s, 2, 1, s->Vz, s->Rz, 0, 0, 0, 0, 24);
}

/* WavesElastic3D::backwardstepStress (waves.cpp:3715-4103): each stress
 * gains dt times its velocity derivatives, with the forward strips. */
void FN(orc_e3d_stress_bw)(FN(orc_e3d_t) *s)
{
    int nx = s->nx, ny = s->ny, nz = s->nz, o = s->order;
    long nxy = (long)nx * ny;
    FN(orc_deriv)(s->Vx, s->df, nx, ny, nz, 0, 0, o, s->dx);
    FN(upd3e_)(s, 0, 0, s->Sxx, 0, 0, 0, 0, 0, 6);
    FN(orc_deriv)(s->Vy, s->df, nx, ny, nz, 1, 0, o, s->dy);
    FN(upd3e_)(s, 1, 0, s->Syy, 0, 0, 0, 0, 0, 18);
    FN(orc_deriv)(s->Vz, s->df, nx, ny, nz, 2, 0, o, s->dz);
    FN(upd3e_)(s, 2, 0, s->Szz, 0, 0, 0, 0, 0, 30);
    if (s->fs) /* plane lpml lies in no z strip */
        for (long p = 0; p < nxy; p++) s->Szz[(long)s->lpml * nxy + p] = 0.0;
    FN(orc_deriv)(s->Vz, s->df, nx, ny, nz, 0, 1, o, s->dx);
    FN(upd3e_)(s, 0, 1, s->Sxz, 0, 0, 0, 0, 0, 8);
    FN(orc_deriv)(s->Vx, s->df, nx, ny, nz, 2, 1, o, s->dz);
    FN(upd3e_)(s, 2, 1, s->Sxz, 0, 0, 0, 0, 0, 32);
    FN(orc_deriv)(s->Vy, s->df, nx, ny, nz, 2, 1, o, s->dz);
    FN(upd3e_)(s, 2, 1, s->Syz, 0, 0, 0, 0, 0, 34);
    FN(orc_deriv)(s->Vz, s->df, nx, ny, nz, 1, 1, o, s->dy);
    FN(upd3e_)(s, 1, 1, s->Syz, 0, 0, 0, 0, 0, 22);
    FN(orc_deriv)(s->Vx, s->df, nx, ny, nz, 1, 1, o, s->dy);
    FN(upd3e_)(s, 1, 1, s->Sxy, 0, 0, 0, 0, 0, 20);
    FN(orc_deriv)(s->Vy, s->df, nx, ny, nz, 0, 1, o, s->dx);
    FN(upd3e_)(s, 0, 1, s->Sxy, 0, 0, 0, 0, 0, 10);
}

/* Explosive source of WavesElastic3D::insertSource (waves.cpp:4137-4160):
 * Sxx, Syy, Szz -= dt*wav*xtri*ytri*ztri at each in-bounds trace's 8 cells. */
void FN(orc_inject3d_explosive)(REAL *Sxx, REAL *Syy, REAL *Szz, int nx, int ny, int lpml, REAL dt, const int *map,
                                const REAL *shift, const REAL *wav, int ntrace, int nt, int it)
{
    long nxy = (long)nx * ny;
    for (int i = 0; i < ntrace; i++) {
        const int *m = map + 3 * i;
        const REAL *sh = shift + 3 * i;
        if (m[0] < 0 || m[1] < 0 || m[2] < 0) continue;
        REAL xt[2] = {1 - sh[0], sh[0]}, yt[2] = {1 - sh[1], sh[1]}, zt[2] = {1 - sh[2], sh[2]};
        for (int i1 = 0; i1 < 2; i1++)
            for (int i2 = 0; i2 < 2; i2++)
                for (int i3 = 0; i3 < 2; i3++) {
                    long p = (long)(lpml + m[2] + i1) * nxy + (long)(lpml + m[1] + i2) * nx + (lpml + m[0] + i3);
                    REAL v = dt * wav[(long)i * nt + it] * xt[i3] * yt[i2] * zt[i1];
                    Sxx[p] -= v;
                    Syy[p] -= v;
                    Szz[p] -= v;
                }
    }
}

/* Pressure gather of WavesElastic3D::recordData (waves.cpp:4283-4311): per
 * trace factor = (3 rho vp^2 - 4 rho vs^2) / 3 at the trace's model cell
 * (in double, rounded to REAL), then data += (1/3) w F / factor for F =
 * Sxx, Syy, Szz at each corner, in double. */
void FN(orc_record3d_pressure_el)(const REAL *Sxx, const REAL *Syy, const REAL *Szz, int nx, int ny, int lpml,
                                  const REAL *vp, const REAL *vs, const REAL *rho, int mnx, int mny, const int *map,
                                  const REAL *shift, REAL *data, int ntrace, int nt, int it)
{
    long nxy = (long)nx * ny;
    const REAL *F[3] = {Sxx, Syy, Szz};
    for (int i = 0; i < ntrace; i++) {
        const int *m = map + 3 * i;
        const REAL *sh = shift + 3 * i;
        if (m[0] < 0 || m[1] < 0 || m[2] < 0) continue;
        long c = ((long)m[2] * mny + m[1]) * mnx + m[0];
        REAL factor = (REAL)((3. * (double)rho[c] * (double)vp[c] * (double)vp[c] -
                              4. * (double)rho[c] * (double)vs[c] * (double)vs[c]) / 3.);
        REAL xt[2] = {1 - sh[0], sh[0]}, yt[2] = {1 - sh[1], sh[1]}, zt[2] = {1 - sh[2], sh[2]};
        long q = (long)i * nt + it;
        for (int i1 = 0; i1 < 2; i1++)
            for (int i2 = 0; i2 < 2; i2++)
                for (int i3 = 0; i3 < 2; i3++) {
                    long p = (long)(lpml + m[2] + i1) * nxy + (long)(lpml + m[1] + i2) * nx + (lpml + m[0] + i3);
                    for (int k = 0; k < 3; k++)
                        data[q] = (REAL)((double)data[q] + (1. / 3.) * (double)xt[i3] * (double)yt[i2] *
                                                               (double)zt[i1] * (double)F[k][p] / (double)factor);
                }
    }
}

/* ---------------------------------------------------------------------- */
/* Model staggering (model.cpp:93-239, 1350-1411, 1965-2040, 2720-2801)    */
/* ---------------------------------------------------------------------- */

/* Edge-replication padding of an nx*nz model into (nx+2L)*(nz+2L). */
static void FN(pad2d_)(REAL *out, const REAL *in, int nx, int nz, int L)
{
    int nxp = nx + 2 * L, nzp = nz + 2 * L;
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++) out[(long)(iz + L) * nxp + ix + L] = in[(long)iz * nx + ix];
    for (int ix = 0; ix < L; ix++)
        for (int iz = 0; iz < nzp; iz++) {
            out[(long)iz * nxp + ix] = out[(long)iz * nxp + L];
            out[(long)iz * nxp + nxp - L + ix] = out[(long)iz * nxp + nxp - L - 1];
        }
    for (int ix = 0; ix < nxp; ix++)
        for (int iz = 0; iz < L; iz++) {
            out[(long)iz * nxp + ix] = out[(long)L * nxp + ix];
            out[(long)(nzp - L + iz) * nxp + ix] = out[(long)(nzp - L - 1) * nxp + ix];
        }
}

/* Arithmetic-mean staggering along one axis in place; last plane copies
 * its neighbour (model.cpp:185-239). */
static void FN(stagger_)(REAL *m, int nx, int ny, int nz, int axis)
{
    long nxy = (long)nx * ny;
    long s = axis == 0 ? 1 : (axis == 1 ? nx : nxy);
    int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
    for (int iz = 0; iz < (axis == 2 ? 1 : nz); iz++)
        for (int iy = 0; iy < (axis == 1 ? 1 : ny); iy++)
            for (int ix = 0; ix < (axis == 0 ? 1 : nx); ix++) {
                long b = iz * nxy + (long)iy * nx + ix;
                for (int i = 0; i < n - 1; i++) {
                    REAL avg = (REAL)((double)(m[b + (i + 1) * s] + m[b + i * s]) / 2.0);
                    m[b + i * s] = avg;
                }
                m[b + (long)(n - 1) * s] = m[b + (long)(n - 2) * s];
            }
}

static int FN(invert_)(REAL *m, long n)
{
    for (long p = 0; p < n; p++) {
        if (m[p] == 0.0) return -1;
        m[p] = (REAL)(1.0 / (double)m[p]);
    }
    return 0;
}

/* ModelAcoustic2D::staggerModels: L=rho*vp^2, Rx/Rz = 1/avg(rho).
 * Outputs are (nx+2L)*(nz+2L).  Returns -1 on zero density. */
int FN(orc_stagger_a2d)(const REAL *vp, const REAL *rho, int nx, int nz, int lpml, int fs,
                        REAL *L, REAL *Rx, REAL *Rz)
{
    int nxp = nx + 2 * lpml, nzp = nz + 2 * lpml;
    long n = (long)nxp * nzp;
    FN(pad2d_)(Rx, rho, nx, nz, lpml);
    FN(pad2d_)(Rz, rho, nx, nz, lpml);
    FN(pad2d_)(L, vp, nx, nz, lpml);
    for (long p = 0; p < n; p++) {
        REAL v = L[p], r = Rx[p];
        L[p] = r * v * v;
    }
    FN(stagger_)(Rx, nxp, 1, nzp, 0);
    FN(stagger_)(Rz, nxp, 1, nzp, 2);
    if (FN(invert_)(Rx, n) || FN(invert_)(Rz, n)) return -1;
    if (fs)
        for (int ix = 0; ix < nxp; ix++) {
            Rx[(long)lpml * nxp + ix] *= 2.0;
            L[(long)lpml * nxp + ix] *= 0.0;
        }
    return 0;
}

/* ModelElastic2D::staggerModels (model.cpp:2720-2801). */
int FN(orc_stagger_e2d)(const REAL *vp, const REAL *vs, const REAL *rho, int nx, int nz,
                        int lpml, int fs, REAL *L, REAL *L2M, REAL *M, REAL *Rx, REAL *Rz)
{
    int nxp = nx + 2 * lpml, nzp = nz + 2 * lpml;
    long n = (long)nxp * nzp;
    FN(pad2d_)(Rx, rho, nx, nz, lpml);
    FN(pad2d_)(Rz, rho, nx, nz, lpml);
    FN(pad2d_)(L, vp, nx, nz, lpml);
    FN(pad2d_)(M, vs, nx, nz, lpml);
    for (long p = 0; p < n; p++) {
        REAL v = L[p], w = M[p], r = Rx[p];
        L2M[p] = r * v * v;
        M[p] = r * w * w;
        L[p] = (REAL)((double)(r * v * v) - 2.0 * (double)r * (double)w * (double)w);
    }
    if (fs)
        for (int ix = 0; ix < nxp; ix++) {
            L[(long)lpml * nxp + ix] = 0.0;
            L2M[(long)lpml * nxp + ix] = M[(long)lpml * nxp + ix];
        }
    FN(stagger_)(Rx, nxp, 1, nzp, 0);
    FN(stagger_)(Rz, nxp, 1, nzp, 2);
    FN(stagger_)(M, nxp, 1, nzp, 2);
    FN(stagger_)(M, nxp, 1, nzp, 0);
    if (FN(invert_)(Rx, n) || FN(invert_)(Rz, n)) return -1;
    if (fs)
        for (int ix = 0; ix < nxp; ix++) Rx[(long)lpml * nxp + ix] *= 2.0;
    return 0;
}

/* Model::padmodel3d (model.cpp:130-182): the model in the interior, every
 * padding cell a copy of the nearest model cell (edge replication). */
static void FN(pad3d_)(REAL *out, const REAL *in, int nx, int ny, int nz, int L)
{
    int nxp = nx + 2 * L, nyp = ny + 2 * L, nzp = nz + 2 * L;
    for (int iz = 0; iz < nzp; iz++)
        for (int iy = 0; iy < nyp; iy++)
            for (int ix = 0; ix < nxp; ix++) {
                int jx = ix - L < 0 ? 0 : (ix - L >= nx ? nx - 1 : ix - L);
                int jy = iy - L < 0 ? 0 : (iy - L >= ny ? ny - 1 : iy - L);
                int jz = iz - L < 0 ? 0 : (iz - L >= nz ? nz - 1 : iz - L);
                out[((long)iz * nyp + iy) * nxp + ix] = in[((long)jz * ny + jy) * nx + jx];
            }
}

/* ModelElastic3D::staggerModels (model.cpp:3371-3473): pad (padmodel3d),
 * L2M = rho vp^2, L = rho vp^2 - 2 rho vs^2 (in double), M = rho vs^2;
 * free surface plane lpml: L = 0, L2M = M; Rx/Ry/Rz staggered along x/y/z,
 * M_xz along x then z, M_yz y then z, M_xy x then y; densities inverted;
 * free surface: Rx, Ry doubled on plane lpml.  out: L, L2M, Mxz, Myz, Mxy,
 * Rx, Ry, Rz, each (nx+2L)(ny+2L)(nz+2L). */
int FN(orc_stagger_e3d)(const REAL *vp, const REAL *vs, const REAL *rho, int nx, int ny, int nz, int lpml, int fs,
                        REAL *L, REAL *L2M, REAL *Mxz, REAL *Myz, REAL *Mxy, REAL *Rx, REAL *Ry, REAL *Rz)
{
    int nxp = nx + 2 * lpml, nyp = ny + 2 * lpml, nzp = nz + 2 * lpml;
    long nxy = (long)nxp * nyp, n = nxy * nzp;
    FN(pad3d_)(Rx, rho, nx, ny, nz, lpml);
    FN(pad3d_)(Ry, rho, nx, ny, nz, lpml);
    FN(pad3d_)(Rz, rho, nx, ny, nz, lpml);
    FN(pad3d_)(L, vp, nx, ny, nz, lpml);
    FN(pad3d_)(Mxz, vs, nx, ny, nz, lpml);
    for (long p = 0; p < n; p++) {
        REAL v = L[p], w = Mxz[p], r = Rx[p];
        L2M[p] = r * v * v;
        L[p] = (REAL)((double)(r * v * v) - 2.0 * (double)r * (double)w * (double)w);
        Mxz[p] = r * w * w;
        Myz[p] = r * w * w;
        Mxy[p] = r * w * w;
    }
    if (fs)
        for (long p = 0; p < nxy; p++) {
            L[(long)lpml * nxy + p] = 0.0;
            L2M[(long)lpml * nxy + p] = Mxz[(long)lpml * nxy + p];
        }
    FN(stagger_)(Rx, nxp, nyp, nzp, 0);
    FN(stagger_)(Ry, nxp, nyp, nzp, 1);
    FN(stagger_)(Rz, nxp, nyp, nzp, 2);
    FN(stagger_)(Mxz, nxp, nyp, nzp, 0);
    FN(stagger_)(Mxz, nxp, nyp, nzp, 2);
    FN(stagger_)(Myz, nxp, nyp, nzp, 1);
    FN(stagger_)(Myz, nxp, nyp, nzp, 2);
    FN(stagger_)(Mxy, nxp, nyp, nzp, 0);
    FN(stagger_)(Mxy, nxp, nyp, nzp, 1);
    if (FN(invert_)(Rx, n) || FN(invert_)(Ry, n) || FN(invert_)(Rz, n)) return -1;
    if (fs)
        for (long p = 0; p < nxy; p++) {
            Rx[(long)lpml * nxy + p] *= 2.0;
            Ry[(long)lpml * nxy + p] *= 2.0;
        }
    return 0;
}

/* ModelAcoustic3D::staggerModels (model.cpp:1965-2040): pad in y, x, z
 * order (padmodel3d), L=rho*vp^2, Rx/Ry/Rz staggered and inverted;
 * free surface doubles Rx and Ry and zeroes L on plane z=lpml. */
int FN(orc_stagger_a3d)(const REAL *vp, const REAL *rho, int nx, int ny, int nz, int lpml,
                        int fs, REAL *L, REAL *Rx, REAL *Ry, REAL *Rz)
{
    int nxp = nx + 2 * lpml, nyp = ny + 2 * lpml, nzp = nz + 2 * lpml;
    long nxy = (long)nxp * nyp, n = nxy * nzp;
    REAL *dst[4] = {Rx, Ry, Rz, L};
    const REAL *src[4] = {rho, rho, rho, vp};
    for (int a = 0; a < 4; a++) {
        REAL *o = dst[a];
        const REAL *m = src[a];
        for (int ix = 0; ix < nx; ix++)
            for (int iy = 0; iy < ny; iy++)
                for (int iz = 0; iz < nz; iz++)
                    o[(long)(iz + lpml) * nxy + (long)(iy + lpml) * nxp + ix + lpml] =
                        m[((long)iz * ny + iy) * nx + ix];
        for (int ix = 0; ix < nxp; ix++)
            for (int iy = 0; iy < lpml; iy++)
                for (int iz = 0; iz < nzp; iz++) {
                    o[iz * nxy + (long)iy * nxp + ix] = o[iz * nxy + (long)lpml * nxp + ix];
                    o[iz * nxy + (long)(nyp - lpml + iy) * nxp + ix] = o[iz * nxy + (long)(nyp - lpml - 1) * nxp + ix];
                }
        for (int ix = 0; ix < lpml; ix++)
            for (int iy = 0; iy < nyp; iy++)
                for (int iz = 0; iz < nzp; iz++) {
                    o[iz * nxy + (long)iy * nxp + ix] = o[iz * nxy + (long)iy * nxp + lpml];
                    o[iz * nxy + (long)iy * nxp + nxp - lpml + ix] = o[iz * nxy + (long)iy * nxp + nxp - lpml - 1];
                }
        for (int ix = 0; ix < nxp; ix++)
            for (int iy = 0; iy < nyp; iy++)
                for (int iz = 0; iz < lpml; iz++) {
                    o[iz * nxy + (long)iy * nxp + ix] = o[(long)lpml * nxy + (long)iy * nxp + ix];
                    o[(nzp - lpml + iz) * nxy + (long)iy * nxp + ix] = o[(long)(nzp - lpml - 1) * nxy + (long)iy * nxp + ix];
                }
    }
    for (long p = 0; p < n; p++) {
        REAL v = L[p], r = Rx[p];
        L[p] = r * v * v;
    }
    FN(stagger_)(Rx, nxp, nyp, nzp, 0);
    FN(stagger_)(Ry, nxp, nyp, nzp, 1);
    FN(stagger_)(Rz, nxp, nyp, nzp, 2);
    if (FN(invert_)(Rx, n) || FN(invert_)(Ry, n) || FN(invert_)(Rz, n)) return -1;
    if (fs)
        for (int ix = 0; ix < nxp; ix++)
            for (int iy = 0; iy < nyp; iy++) {
                long p = (long)lpml * nxy + (long)iy * nxp + ix;
                Rx[p] *= 2.0;
                Ry[p] *= 2.0;
                L[p] *= 0.0;
            }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* FwiAcoustic2D imaging condition and misfit (fwi.cpp:511-800)            */
/* ---------------------------------------------------------------------- */

/*
 * crossCorr for one adjoint step, non-domdec form: interior ix,iz in
 * [1,n-1) of the unpadded nx*nz image; wsp is the forward pressure on the
 * padded grid (pad `pads`), wrx/wrz the adjoint velocities (pad `padr`);
 * Vp/Rho unpadded, Rx/Rz padded (stride nx+2*padr).
 */
void FN(orc_a2d_crosscorr)(const REAL *wsp, int pads, const REAL *wrx, const REAL *wrz,
                           int padr, const REAL *Vp, const REAL *Rho, const REAL *Rx,
                           const REAL *Rz, int nx, int nz, REAL dx, REAL dz,
                           REAL *vpgrad, REAL *rhograd)
{
    int nxs = nx + 2 * pads, nxr = nx + 2 * padr;
#define KS(i, j) ((long)(j) * nxs + (i))
#define KR(i, j) ((long)(j) * nxr + (i))
#define KM(i, j) ((long)(j) * nx + (i))
    for (int ix = 1; ix < nx - 1; ix++)
        for (int iz = 1; iz < nz - 1; iz++) {
            REAL vpscale = (REAL)(-2.0 / (double)Vp[KM(ix, iz)]);
            REAL rhoscale1 = (REAL)(-1.0 / (double)Rho[KM(ix, iz)]);
            REAL mrxx = (wrx[KR(ix + padr, iz + padr)] - wrx[KR(ix + padr - 1, iz + padr)]) / dx;
            REAL mrzz = (wrz[KR(ix + padr, iz + padr)] - wrz[KR(ix + padr, iz + padr - 1)]) / dz;
            REAL ws = wsp[KS(ix + pads, iz + pads)];
            vpgrad[KM(ix, iz)] -= vpscale * ws * (mrxx + mrzz);
            rhograd[KM(ix, iz)] -= rhoscale1 * ws * (mrxx + mrzz);
            REAL uderx, uderz;
            uderx = (REAL)(0.5 * (double)wrx[KR(ix + padr, iz + padr)] * (double)Rx[KR(ix + padr, iz + padr)] *
                           (double)(wsp[KS(ix + pads + 1, iz + pads)] - ws) / (double)dx);
            uderx = (REAL)((double)uderx + 0.5 * (double)wrx[KR(ix + padr - 1, iz + padr)] *
                           (double)Rx[KR(ix + padr - 1, iz + padr)] *
                           (double)(ws - wsp[KS(ix + pads - 1, iz + pads)]) / (double)dx);
            uderz = (REAL)(0.5 * (double)wrz[KR(ix + padr, iz + padr)] * (double)Rz[KR(ix + padr, iz + padr)] *
                           (double)(wsp[KS(ix + pads, iz + pads + 1)] - ws) / (double)dz);
            uderz = (REAL)((double)uderz + 0.5 * (double)wrz[KR(ix + padr, iz + padr - 1)] *
                           (double)Rz[KR(ix + padr, iz + padr - 1)] *
                           (double)(ws - wsp[KS(ix + pads, iz + pads - 1)]) / (double)dz);
            rhograd[KM(ix, iz)] += (uderx + uderz);
        }
#undef KS
#undef KR
#undef KM
}

/*
 * FwiAcoustic3D::crossCorr (fwi.cpp:1403-1466), non-domdec form, with the
 * reference's loop bounds ix, iy, iz in [1, n) (the upper edge included,
 * unlike 2D) and its index of the second z term: wrz at (ix, iy-1, iz)
 * against Rz at (ix, iy, iz-1) (fwi.cpp:1459).  wsp forward pressure (pad
 * pads), wrx/wry/wrz adjoint velocities and Rx/Ry/Rz (pad padr); Vp/Rho and
 * the images unpadded nx*ny*nz.
 */
void FN(orc_a3d_crosscorr)(const REAL *wsp, int pads, const REAL *wrx, const REAL *wry, const REAL *wrz,
                           int padr, const REAL *Vp, const REAL *Rho, const REAL *Rx, const REAL *Ry,
                           const REAL *Rz, int nx, int ny, int nz, REAL dx, REAL dy, REAL dz,
                           REAL *vpgrad, REAL *rhograd)
{
    const long nxs = nx + 2 * pads, nys = ny + 2 * pads, nxr = nx + 2 * padr, nyr = ny + 2 * padr;
#define KS(i, j, k) ((long)(k) * nxs * nys + (long)(j) * nxs + (i))
#define KR(i, j, k) ((long)(k) * nxr * nyr + (long)(j) * nxr + (i))
#define KM(i, j, k) ((long)(k) * nx * ny + (long)(j) * nx + (i))
    for (int ix = 1; ix < nx; ix++)
        for (int iy = 1; iy < ny; iy++)
            for (int iz = 1; iz < nz; iz++) {
                const int xs = ix + pads, ys = iy + pads, zs = iz + pads;
                const int xr = ix + padr, yr = iy + padr, zr = iz + padr;
                REAL vpscale = (REAL)(-2.0 / (double)Vp[KM(ix, iy, iz)]);
                REAL rhoscale1 = (REAL)(-1.0 / (double)Rho[KM(ix, iy, iz)]);
                REAL mrxx = (wrx[KR(xr, yr, zr)] - wrx[KR(xr - 1, yr, zr)]) / dx;
                REAL mryy = (wry[KR(xr, yr, zr)] - wry[KR(xr, yr - 1, zr)]) / dy;
                REAL mrzz = (wrz[KR(xr, yr, zr)] - wrz[KR(xr, yr, zr - 1)]) / dz;
                REAL ws = wsp[KS(xs, ys, zs)];
                vpgrad[KM(ix, iy, iz)] -= vpscale * ws * (mrxx + mryy + mrzz);
                rhograd[KM(ix, iy, iz)] -= rhoscale1 * ws * (mrxx + mryy + mrzz);
                REAL uderx, udery, uderz;
                uderx = (REAL)(0.5 * (double)wrx[KR(xr, yr, zr)] * (double)Rx[KR(xr, yr, zr)] *
                               (double)(wsp[KS(xs + 1, ys, zs)] - ws) / (double)dx);
                uderx = (REAL)((double)uderx + 0.5 * (double)wrx[KR(xr - 1, yr, zr)] * (double)Rx[KR(xr - 1, yr, zr)] *
                               (double)(ws - wsp[KS(xs - 1, ys, zs)]) / (double)dx);
                udery = (REAL)(0.5 * (double)wry[KR(xr, yr, zr)] * (double)Ry[KR(xr, yr, zr)] *
                               (double)(wsp[KS(xs, ys + 1, zs)] - ws) / (double)dy);
                udery = (REAL)((double)udery + 0.5 * (double)wry[KR(xr, yr - 1, zr)] * (double)Ry[KR(xr, yr - 1, zr)] *
                               (double)(ws - wsp[KS(xs, ys - 1, zs)]) / (double)dy);
                uderz = (REAL)(0.5 * (double)wrz[KR(xr, yr, zr)] * (double)Rz[KR(xr, yr, zr)] *
                               (double)(wsp[KS(xs, ys, zs + 1)] - ws) / (double)dz);
                uderz = (REAL)((double)uderz + 0.5 * (double)wrz[KR(xr, yr - 1, zr)] * (double)Rz[KR(xr, yr, zr - 1)] *
                               (double)(ws - wsp[KS(xs, ys, zs - 1)]) / (double)dz);
                rhograd[KM(ix, iy, iz)] += (uderx + udery + uderz);
            }
#undef KS
#undef KR
#undef KM
}

/*
 * FwiElastic2D::crossCorr(waves_fw, waves_bw, model, it) (fwi.cpp:2591-2775),
 * non-domdec form: interior ix,iz in [1,n-1) of the unpadded nx*nz images.
 * Forward fields ws* and adjoint fields wr*, rs* on padded grids (both pad
 * lpml), Rx/Rz padded.  lam/mu/rho accumulate the kernels before scaleGrad
 * (the reference's vpgrad/vsgrad/rhograd images at that point).  When
 * wavgrad != NULL, the adjoint field of `wfield` (0 PRESSURE: Sxx+Szz,
 * 1 VX: Vx, 3 VZ: Vz) is gathered bilinearly at the source traces whose map
 * falls on this interior cell, into wavgrad[it + nt*i].
 */
void FN(orc_e2d_crosscorr)(const REAL *wsx, const REAL *wsz, const REAL *wsxx, const REAL *wszz,
                           const REAL *wsxz, const REAL *wrx, const REAL *wrz, const REAL *rsxx,
                           const REAL *rszz, const REAL *rsxz, const REAL *Rx, const REAL *Rz,
                           int lpml, int nx, int nz, REAL dx, REAL dz, REAL *lam, REAL *mu,
                           REAL *rho, REAL *wavgrad, int wfield, const int *smap, const REAL *sshift,
                           int ntrace, int nt, int it)
{
    const int pads = lpml, padr = lpml;
    const int nxs = nx + 2 * pads, nxr = nx + 2 * padr;
#define KS(i, j) ((long)(j) * nxs + (i))
#define KR(i, j) ((long)(j) * nxr + (i))
#define KI(i, j) ((long)(j) * nx + (i))
    for (int ix = 1; ix < nx - 1; ix++)
        for (int iz = 1; iz < nz - 1; iz++) {
            const int xs = ix + pads, zs = iz + pads, xr = ix + padr, zr = iz + padr;
            REAL msxx = (wsx[KS(xs, zs)] - wsx[KS(xs - 1, zs)]) / dx;
            REAL mszz = (wsz[KS(xs, zs)] - wsz[KS(xs, zs - 1)]) / dz;
            lam[KI(ix, iz)] -= (msxx + mszz) * (rsxx[KR(xr, zr)] + rszz[KR(xr, zr)]);
            REAL msxz1 = (wsx[KS(xs - 1, zs)] - wsx[KS(xs - 1, zs - 1)]) / dz +
                         (wsz[KS(xs, zs - 1)] - wsz[KS(xs - 1, zs - 1)]) / dx;
            REAL msxz2 = (wsx[KS(xs, zs)] - wsx[KS(xs, zs - 1)]) / dz +
                         (wsz[KS(xs + 1, zs - 1)] - wsz[KS(xs, zs - 1)]) / dx;
            REAL msxz3 = (wsx[KS(xs - 1, zs + 1)] - wsx[KS(xs - 1, zs)]) / dz +
                         (wsz[KS(xs, zs)] - wsz[KS(xs - 1, zs)]) / dx;
            REAL msxz4 = (wsx[KS(xs, zs + 1)] - wsx[KS(xs, zs)]) / dz +
                         (wsz[KS(xs + 1, zs)] - wsz[KS(xs, zs)]) / dx;
            REAL mrxz1 = rsxz[KR(xr - 1, zr - 1)];
            REAL mrxz2 = rsxz[KR(xr, zr - 1)];
            REAL mrxz3 = rsxz[KR(xr - 1, zr)];
            REAL mrxz4 = rsxz[KR(xr, zr)];
            REAL f1 = msxx * rszz[KR(xr, zr)] + mszz * rsxx[KR(xr, zr)];
            REAL f2 = msxz1 * mrxz1 + msxz2 * mrxz2 + msxz3 * mrxz3 + msxz4 * mrxz4;
            mu[KI(ix, iz)] = (REAL)((double)mu[KI(ix, iz)] + (2.0 * (double)f1 - 0.25 * (double)f2));
            if (wavgrad) {
                for (int i = 0; i < ntrace; i++) {
                    if (smap[2 * i] != ix || smap[2 * i + 1] != iz) continue;
                    REAL xtri[2], ytri[2];
                    xtri[0] = 1 - sshift[2 * i];
                    xtri[1] = sshift[2 * i];
                    ytri[0] = 1 - sshift[2 * i + 1];
                    ytri[1] = sshift[2 * i + 1];
                    for (int i1 = 0; i1 < 2; i1++)
                        for (int i2 = 0; i2 < 2; i2++) {
                            long q = KR(xr + i2, zr + i1);
                            REAL v = wfield == 0 ? rsxx[q] + rszz[q] : (wfield == 1 ? wrx[q] : wrz[q]);
                            wavgrad[(long)i * nt + it] += xtri[i2] * ytri[i1] * v;
                        }
                }
            }
            /* rho kernel: double temporaries stored to T after each line */
            REAL uderx, uderz;
            uderx = (REAL)(0.5 * (double)wrx[KS(xs, zs)] * (double)Rx[KR(xr, zr)] *
                           (double)(wsxx[KR(xr + 1, zr)] - wsxx[KR(xr, zr)]) / (double)dx);
            uderx = (REAL)((double)uderx + 0.5 * (double)wrx[KS(xs - 1, zs)] * (double)Rx[KR(xr - 1, zr)] *
                           (double)(wsxx[KR(xr, zr)] - wsxx[KR(xr - 1, zr)]) / (double)dx);
            uderx = (REAL)((double)uderx + 0.5 * (double)wrx[KS(xs, zs)] * (double)Rx[KR(xr, zr)] *
                           (double)(wsxz[KR(xr, zr)] - wsxz[KR(xr, zr - 1)]) / (double)dz);
            uderx = (REAL)((double)uderx + 0.5 * (double)wrx[KS(xs - 1, zs)] * (double)Rx[KR(xr - 1, zr)] *
                           (double)(wsxz[KR(xr - 1, zr)] - wsxz[KR(xr - 1, zr - 1)]) / (double)dz);
            uderz = (REAL)(0.5 * (double)wrz[KS(xs, zs)] * (double)Rz[KR(xr, zr)] *
                           (double)(wsxz[KR(xr, zr)] - wsxz[KR(xr - 1, zr)]) / (double)dx);
            uderz = (REAL)((double)uderz + 0.5 * (double)wrz[KS(xs, zs - 1)] * (double)Rz[KR(xr, zr - 1)] *
                           (double)(wsxz[KR(xr, zr - 1)] - wsxz[KR(xr - 1, zr - 1)]) / (double)dx);
            uderz = (REAL)((double)uderz + 0.5 * (double)wrz[KS(xs, zs)] * (double)Rz[KR(xr, zr)] *
                           (double)(wszz[KR(xr, zr + 1)] - wszz[KR(xr, zr)]) / (double)dz);
            uderz = (REAL)((double)uderz + 0.5 * (double)wrz[KS(xs, zs - 1)] * (double)Rz[KR(xr, zr - 1)] *
                           (double)(wszz[KR(xr, zr)] - wszz[KR(xr, zr - 1)]) / (double)dz);
            rho[KI(ix, iz)] -= (uderx + uderz);
        }
#undef KS
#undef KR
#undef KI
}

/* FwiElastic2D::scaleGrad (fwi.cpp:2778-2849): lambda/mu/rho kernels ->
 * Vp/Vs/rho gradients in place, over the whole unpadded model. */
void FN(orc_e2d_scalegrad)(const REAL *Vp, const REAL *Vs, const REAL *Rho, int nx, int nz,
                           REAL *vpgrad, REAL *vsgrad, REAL *rhograd)
{
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++) {
            long k = (long)iz * nx + ix;
            REAL vpscale = (REAL)(2.0 * (double)Rho[k] * (double)Vp[k]);
            REAL vsscale = (REAL)(2.0 * (double)Rho[k] * (double)Vs[k]);
            REAL rhoscale1 = (REAL)((double)(Vp[k] * Vp[k]) - 2.0 * (double)Vs[k] * (double)Vs[k]);
            REAL rhoscale2 = Vs[k] * Vs[k];
            REAL lambda = vpgrad[k], mu = vsgrad[k], rho = rhograd[k];
            vpgrad[k] = vpscale * lambda;
            vsgrad[k] = (REAL)((double)(vsscale * mu) - 2.0 * (double)vsscale * (double)lambda);
            rhograd[k] = rhoscale1 * lambda + rhoscale2 * mu + rho;
        }
}

/* Elastic DIFFERENCE misfit over the P, Ux, Uz records (fwi.cpp:2852-2930):
 * misfit += 0.5*(rp^2 + rx^2 + rz^2) in T, traces whose Ux map is valid. */
REAL FN(orc_misfit_diff3)(const REAL *modp, const REAL *recp, const REAL *modx, const REAL *recx,
                          const REAL *modz, const REAL *recz, const int *valid, int ntr, int nt)
{
    REAL misfit = 0.0;
    for (int i = 0; i < ntr; i++) {
        if (!valid[i]) continue;
        for (int it = 0; it < nt; it++) {
            long q = (long)i * nt + it;
            REAL rp = modp[q] - recp[q], rx = modx[q] - recx[q], rz = modz[q] - recz[q];
            misfit = (REAL)((double)misfit + 0.5 * (double)(rp * rp + rx * rx + rz * rz));
        }
    }
    return misfit;
}

/*
 * Interp::interp (interp.cpp:110-186) for trace-major data: from[ntr][n1]
 * sampled at o1 + i*d1 -> to[ntr][n2] at o2 + j*d2; mode 0 LINEAR (2 taps,
 * interp.cpp:69-76), 1 BSPLINE (4 taps, :79-97), 2 SINC (6 taps, Lanczos-3
 * window, :99-125).  Output samples whose taps leave [0, n1) are zero.  The
 * reference's float / double promotions are kept (T = REAL).
 */
static REAL FN(orc_sinc)(REAL x)
{
    if (x == 0.0) return (REAL)1.0;
    return (REAL)(sin(3.14159265358979323846 * (double)x) / (3.14159265358979323846 * (double)x));
}

void FN(orc_interp)(int mode, const REAL *from, int ntr, int n1, REAL d1, REAL o1, REAL *to, int n2, REAL d2,
                    REAL o2)
{
    const int pad = mode == 0 ? 1 : (mode == 1 ? 2 : 3);
    for (int i = 0; i < ntr; i++)
        for (long j = 0; j < n2; j++) {
            REAL val = o2 + (REAL)j * d2;
            long i0 = (long)((val - o1) / d1) - (pad - 1);
            REAL *out = to + (long)i * n2 + j;
            if (i0 < 0 || (i0 + (2 * pad) - 1) > (n1 - 1)) {
                *out = (REAL)0.0;
                continue;
            }
            REAL x0 = (o1 + (REAL)(i0 + (pad - 1)) * d1);
            REAL d = (val - x0) / d1;
            const REAL *m = from + (long)i * n1 + i0;
            if (mode == 0) {
                *out = (REAL)((1.0 - (double)d) * (double)m[0] + (double)(d * m[1]));
            } else if (mode == 1) {
                REAL xp1 = (REAL)((double)d + 1.0);
                REAL omx = (REAL)(1.0 - (double)d);
                REAL a0 = (REAL)(((-(1.0 / 6.0) * (double)xp1 + 1.0) * (double)xp1 - 2.0) * (double)xp1 + (4.0 / 3.0));
                REAL a1 = (REAL)((0.5 * (double)d - 1.0) * (double)d * (double)d + (2.0 / 3.0));
                REAL a2 = (REAL)((0.5 * (double)omx - 1.0) * (double)omx * (double)omx + (2.0 / 3.0));
                REAL a3 = (REAL)(1.0 - (double)a0 - (double)a1 - (double)a2);
                *out = a0 * m[0] + a1 * m[1] + a2 * m[2] + a3 * m[3];
            } else {
                REAL acc = (REAL)0.0;
                for (int k = 0; k < 6; k++) {
                    REAL s1 = FN(orc_sinc)((REAL)((2.0 - k) + (double)d));
                    REAL s2 = FN(orc_sinc)((REAL)(((2.0 - k) + (double)d) / 3.0));
                    acc = acc + m[k] * s1 * s2;
                }
                *out = acc;
            }
        }
}

/*
 * FwiAcoustic2D::computeMisfit (fwi.cpp:574-660) for type 0 DIFFERENCE
 * (optional data weights: res *= wei) and 1 CORRELATION (per trace
 * -sum(mod*rec)/(|mod||rec|), norms accumulated sample by sample in T, sqrt
 * rounded to T, zero norms replaced by 1); traces with valid maps only;
 * misfit accumulated in T.  wei may be NULL.
 */
/* ---------------------------------------------------------------------- */
/* ADAPTIVE misfits (fwi.cpp:152-186 stoep, 188-378 convolve, 385-395 xcor, */
/* 397-420 gauss / linear; misfit 647-736, residuals 851-1014)              */
/* ---------------------------------------------------------------------- */

/* z = x convolved with y on the index ranges [ifx, ifx+lx), [ify, ify+ly),
 * output [ifz, ifz+lz): Hale's pairwise scheme as in Fwi::convolve -- zero
 * off the left, the rolling-on outputs two at a time walking j down, the
 * middle outputs two at a time walking j up (a leftover term when the
 * overlap is odd), the rolling-off outputs two at a time, zero off the right;
 * single outputs where a phase has an odd count.  Every sum accumulates its
 * terms in that order, so float results equal the reference's.  Reads the
 * scheme makes past an array end and never uses are not made here. */
static void FN(conv_)(int lx, int ifx, const REAL *x, int ly, int ify, const REAL *y, int lz, int ifz, REAL *z)
{
    int ilx = ifx + lx - 1, ily = ify + ly - 1, ilz = ifz + lz - 1;
    int i, j, ilow, ihigh, jlow, jhigh, t;
    REAL sa, sb, xa, xb, ya, yb;
    if (lx > ly) {
        const REAL *tp;
        t = ifx; ifx = ify; ify = t;
        t = ilx; ilx = ily; ily = t;
        t = lx; lx = ly; ly = t;
        tp = x; x = y; y = tp;
    }
    x -= ifx;
    y -= ify;
    z -= ifz;
    /* off left */
    ilow = ifz;
    ihigh = ify + ifx - 1;
    if (ihigh > ilz) ihigh = ilz;
    for (i = ilow; i <= ihigh; ++i) z[i] = 0.0;
    /* rolling on: one single output first if the overlap count is even */
    if (i < ify + ilx && i <= ilz) {
        jlow = ifx;
        jhigh = i - ify;
        if ((jhigh - jlow) % 2) {
            sa = 0.0;
            for (j = jlow; j <= jhigh; ++j) sa += x[j] * y[i - j];
            z[i++] = sa;
        }
    }
    ilow = i;
    ihigh = ilx + ify - 1;
    if (ihigh > ilz) ihigh = ilz;
    jlow = ifx;
    jhigh = ilow - ify;
    for (i = ilow; i < ihigh; i += 2, jhigh += 2) {
        sa = sb = 0.0;
        xb = x[jhigh + 1];
        yb = 0.0;
        for (j = jhigh; j >= jlow; j -= 2) {
            sa += xb * yb;
            ya = y[i - j];
            sb += xb * ya;
            xa = x[j];
            sa += xa * ya;
            yb = y[i + 1 - j];
            sb += xa * yb;
            if (j - 1 >= jlow) xb = x[j - 1];
        }
        z[i] = sa;
        z[i + 1] = sb;
    }
    if (i == ihigh) {
        jlow = ifx;
        jhigh = i - ify;
        sa = 0.0;
        for (j = jlow; j <= jhigh; ++j) sa += x[j] * y[i - j];
        z[i++] = sa;
    }
    /* middle: full overlap */
    ilow = i;
    ihigh = ily + ifx;
    if (ihigh > ilz) ihigh = ilz;
    jlow = ifx;
    jhigh = ilx;
    for (i = ilow; i < ihigh; i += 2) {
        sa = sb = 0.0;
        yb = y[i + 1 - jlow];
        xa = x[jlow];
        for (j = jlow; j < jhigh; j += 2) {
            sb += xa * yb;
            ya = y[i - j];
            sa += xa * ya;
            xb = x[j + 1];
            sb += xb * ya;
            yb = y[i - 1 - j];
            sa += xb * yb;
            if (j + 2 <= jhigh) xa = x[j + 2];
        }
        if ((jhigh - jlow) % 2) {
            z[i] = sa;
            z[i + 1] = sb;
        } else {
            z[i] = sa + x[jhigh] * y[i - jhigh];
            z[i + 1] = sb + x[jhigh] * y[i + 1 - jhigh];
        }
    }
    if (i == ihigh) {
        sa = 0.0;
        for (j = jlow; j <= jhigh; ++j) sa += x[j] * y[i - j];
        z[i++] = sa;
    }
    /* rolling off: one single output first if the overlap count is odd */
    if (i <= ily + ilx && i <= ilz) {
        jlow = i - ily;
        jhigh = ilx;
        if (!((jhigh - jlow) % 2)) {
            sa = 0.0;
            for (j = jlow; j <= jhigh; ++j) sa += x[j] * y[i - j];
            z[i++] = sa;
        }
    }
    ilow = i;
    ihigh = ily + ilx;
    if (ihigh > ilz) ihigh = ilz;
    jlow = ilow - ily;
    jhigh = ilx - 2;
    for (i = ilow; i < ihigh; i += 2, jlow += 2) {
        sa = sb = 0.0;
        xa = x[jlow];
        yb = 0.0;
        for (j = jlow; j < jhigh; j += 2) {
            sb += xa * yb;
            ya = y[i - j];
            sa += xa * ya;
            xb = x[j + 1];
            sb += xb * ya;
            yb = y[i - 1 - j];
            sa += xb * yb;
            xa = x[j + 2];
        }
        sb += xa * yb;
        ya = y[i - j];
        sa += xa * ya;
        xb = x[j + 1];
        sb += xb * ya;
        yb = y[i - 1 - j];
        sa += xb * yb;
        z[i] = sa;
        z[i + 1] = sb;
    }
    if (i == ihigh) {
        jlow = i - ily;
        jhigh = ilx;
        sa = 0.0;
        for (j = jlow; j <= jhigh; ++j) sa += x[j] * y[i - j];
        z[i++] = sa;
    }
    /* off right */
    for (; i <= ilz; ++i) z[i] = 0.0;
}

/* z = x cross-correlated with y: x reversed, then convolved (xr: lx scratch). */
static void FN(xcor_)(int lx, int ifx, const REAL *x, int ly, int ify, const REAL *y, int lz, int ifz, REAL *z,
                      REAL *xr)
{
    for (int i = 0, j = lx - 1; i < lx; ++i, --j) xr[i] = x[j];
    FN(conv_)(lx, 1 - ifx - lx, xr, ly, ify, y, lz, ifz, z);
}

/* Levinson solve of the symmetric Toeplitz system R f = g (R's first row r),
 * a: scratch (Claerbout's recursion as in Fwi::stoep). */
static void FN(stoep_)(int n, const REAL *r, const REAL *g, REAL *f, REAL *a)
{
    REAL v, e, c, w, bot;
    if (r[0] == 0.0) return;
    a[0] = 1.0;
    v = r[0];
    f[0] = g[0] / r[0];
    for (int j = 1; j < n; j++) {
        int i;
        a[j] = 0.0;
        f[j] = 0.0;
        for (i = 0, e = 0.0; i < j; i++) e += a[i] * r[j - i];
        c = e / v;
        v -= c * e;
        for (i = 0; i <= j / 2; i++) {
            bot = a[j - i] - c * a[i];
            a[i] -= c * a[j - i];
            a[j - i] = bot;
        }
        for (i = 0, w = 0.0; i < j; i++) w += f[i] * r[j - i];
        c = (w - g[j]) / v;
        for (i = 0; i <= j; i++) f[i] -= c * a[j - i];
    }
}

/* Exported for the closed-form checks of the tests. */
void FN(orc_conv)(int lx, int ifx, const REAL *x, int ly, int ify, const REAL *y, int lz, int ifz, REAL *z)
{
    FN(conv_)(lx, ifx, x, ly, ify, y, lz, ifz, z);
}
void FN(orc_stoep)(int n, const REAL *r, const REAL *g, REAL *f, REAL *a) { FN(stoep_)(n, r, g, f, a); }

/* Time weight of the ADAPTIVE misfits: type 2 Gaussian exp(-it^2 / (2 s^2)),
 * s = (T)(0.01 nt); type 3 linear min(it, 500) (fwi.cpp:397-420, fwi.h:28-31). */
static REAL FN(adaptive_h_)(int type, int it, int nt)
{
    if (type == 2) {
        REAL sd = (REAL)(0.01 * nt);
        if (sd == 0.0) return 0;
        return (REAL)exp(-1.0 * (it * it) / (2.0 * (sd * sd)));
    }
    REAL sd = (REAL)500;
    if (it < sd) return (REAL)it;
    return sd;
}

/* Shaping filter of one trace (first half of both ADAPTIVE cases):
 * autocorr of the weighted record, crosscorr record x model, whitening by
 * 1 + 0.01 of the zero lag, shaper = R^-1 crosscorr; returns its squared
 * norm (1 when zero).  w: 6*nt scratch (modwei, recwei, autocorr,
 * crosscorr, spiker, reversed copy), shaper: nt. */
static REAL FN(adaptive_shaper_)(const REAL *m, const REAL *r, const REAL *wt, int nt, REAL *w, REAL *shaper)
{
    REAL *modwei = w, *recwei = w + nt, *autoc = w + 2 * nt, *cross = w + 3 * nt, *spiker = w + 4 * nt,
         *xr = w + 5 * nt;
    for (int it = 0; it < nt; it++) {
        modwei[it] = wt ? m[it] * wt[it] : m[it];
        recwei[it] = wt ? r[it] * wt[it] : r[it];
    }
    FN(xcor_)(nt, 0, recwei, nt, 0, recwei, nt, 0, autoc, xr);
    FN(xcor_)(nt, 0, recwei, nt, 0, modwei, nt, 0, cross, xr);
    if (autoc[0] == 0.0) autoc[0] = 1.0;
    autoc[0] = (REAL)((double)autoc[0] * (1.0 + 0.01));
    for (int it = 0; it < nt; it++) shaper[it] = 0.0;
    FN(stoep_)(nt, autoc, cross, shaper, spiker);
    REAL norm = 0.0;
    for (int it = 0; it < nt; it++) norm += shaper[it] * shaper[it];
    if (norm == 0.0) norm = 1.0;
    return norm;
}

/* Residual of one trace for ADAPTIVE_GAUSS (type 2) / ADAPTIVE_LINEAR (3)
 * (fwi.cpp:851-1014, per component fwi.cpp:3440-3595): the shaper's own
 * misfit, rs = sgn (H^2 - 2 misfit) shaper / norm, solved again with R and
 * convolved with the weighted record, then weighted.  w: 6*nt, shaper: nt. */
static void FN(adaptive_res_)(int type, const REAL *m, const REAL *r, const REAL *wt, REAL *rs, int nt, REAL *w,
                              REAL *shaper, int scaling)
{
    REAL norm = FN(adaptive_shaper_)(m, r, wt, nt, w, shaper);
    REAL misfit = 0.0;
    for (int it = 0; it < nt; it++) {
        REAL h = FN(adaptive_h_)(type, it, nt);
        misfit = (REAL)((double)misfit + 0.5 * (double)(h * shaper[it]) * (double)(h * shaper[it]) / (double)norm);
    }
    for (int it = 0; it < nt; it++) {
        REAL h = FN(adaptive_h_)(type, it, nt);
        double sgn = type == 2 ? -1.0 * scaling : (double)scaling;
        rs[it] = (REAL)(sgn * ((double)(h * h) - 2.0 * (double)misfit) * (double)shaper[it] / (double)norm);
    }
    /* solve again with the residual as right-hand side, then convolve with
     * the weighted record */
    FN(stoep_)(nt, w + 2 * nt, rs, shaper, w + 4 * nt);
    FN(conv_)(nt, 0, shaper, nt, 0, w + nt, nt, 0, rs);
    if (wt)
        for (int it = 0; it < nt; it++) rs[it] *= wt[it];
}

REAL FN(orc_misfit_a)(int type, const REAL *mod, const REAL *rec, const REAL *wei, const int *valid, int ntr,
                      int nt)
{
    if (type == 2 || type == 3) {
        REAL misfit = 0.0;
        REAL *w = (REAL *)malloc(sizeof(REAL) * 7 * (nt > 0 ? nt : 1)), *shaper = w + 6 * nt;
        for (int i = 0; i < ntr; i++) {
            if (!valid[i]) continue;
            const REAL *m = mod + (long)i * nt, *r = rec + (long)i * nt, *wt = wei ? wei + (long)i * nt : NULL;
            REAL norm = FN(adaptive_shaper_)(m, r, wt, nt, w, shaper);
            for (int it = 0; it < nt; it++) {
                REAL res = FN(adaptive_h_)(type, it, nt) * shaper[it];
                if (type == 2)
                    misfit = (REAL)((double)misfit - 0.5 * (double)res * (double)res / (double)norm);
                else
                    misfit = (REAL)((double)misfit + 0.5 * (double)res * (double)res / (double)norm);
            }
        }
        free(w);
        return misfit;
    }
    REAL misfit = 0.0;
    for (int i = 0; i < ntr; i++) {
        if (!valid[i]) continue;
        const REAL *m = mod + (long)i * nt, *r = rec + (long)i * nt, *w = wei ? wei + (long)i * nt : NULL;
        if (type == 1) {
            REAL n1 = 0.0, n2 = 0.0;
            for (int it = 0; it < nt; it++) {
                n1 += m[it] * m[it];
                n2 += r[it] * r[it];
            }
            n1 = (REAL)sqrt((double)n1);
            n2 = (REAL)sqrt((double)n2);
            if (n1 == 0) n1 = 1.0;
            if (n2 == 0) n2 = 1.0;
            for (int it = 0; it < nt; it++) {
                REAL res = (REAL)((-1.0) * (double)(m[it] * r[it] / (n1 * n2)));
                if (w) res *= w[it];
                misfit += res;
            }
        } else {
            for (int it = 0; it < nt; it++) {
                REAL res = m[it] - r[it];
                if (w) res *= w[it];
                misfit = (REAL)((double)misfit + 0.5 * (double)res * (double)res);
            }
        }
    }
    return misfit;
}

static int FN(orc_cmp)(const void *a, const void *b)
{
    REAL x = *(const REAL *)a, y = *(const REAL *)b;
    return (x > y) - (x < y);
}

/*
 * FwiAcoustic2D::computeResiduals (fwi.cpp:803-880) and
 * FwiAcoustic3D::computeResiduals (fwi.cpp:1651-1853): type 0 DIFFERENCE
 * res = scaling*(mod-rec) (*= wei^2); type 1 CORRELATION
 * res = -scaling*(rec/(n1 n2) - mod/(n1 n1) * (sum(mod*rec)/(n1 n2)))
 * (*= wei^2) on every trace; types 2/3 the ADAPTIVE residuals; any other
 * type the default branch, DIFFERENCE.  With clip (the 2D class; the 3D
 * class has no such step) every |res| of types 1-3 at or above the PCLIP =
 * 97th percentile (element (int)(97*nt*ntr/100) of the sorted magnitudes)
 * is then set to zero.
 */
void FN(orc_residual_a)(int type, const REAL *mod, const REAL *rec, const REAL *wei, REAL *res, int ntr, int nt,
                        int scaling, int clip)
{
    const long n = (long)ntr * nt;
    if (type < 1 || type > 3) {
        for (long p = 0; p < n; p++) {
            res[p] = scaling * (mod[p] - rec[p]);
            if (wei) res[p] *= wei[p] * wei[p];
        }
        return;
    }
    if (type == 2 || type == 3) {
        REAL *w = (REAL *)malloc(sizeof(REAL) * 7 * (nt > 0 ? nt : 1)), *shaper = w + 6 * nt;
        for (int i = 0; i < ntr; i++)
            FN(adaptive_res_)(type, mod + (long)i * nt, rec + (long)i * nt, wei ? wei + (long)i * nt : NULL,
                              res + (long)i * nt, nt, w, shaper, scaling);
        free(w);
    } else {
    for (int i = 0; i < ntr; i++) {
        const REAL *m = mod + (long)i * nt, *r = rec + (long)i * nt;
        REAL n1 = 0.0, n2 = 0.0, n3 = 0.0;
        for (int it = 0; it < nt; it++) {
            n1 += m[it] * m[it];
            n2 += r[it] * r[it];
            n3 += m[it] * r[it];
        }
        n1 = (REAL)sqrt((double)n1);
        n2 = (REAL)sqrt((double)n2);
        if (n1 == 0) n1 = 1.0;
        if (n2 == 0) n2 = 1.0;
        n3 /= (n1 * n2);
        for (int it = 0; it < nt; it++) {
            long p = (long)i * nt + it;
            res[p] = (REAL)((-1.0) * scaling * (double)((r[it] / (n1 * n2)) - (m[it] / (n1 * n1)) * n3));
            if (wei) res[p] *= wei[p] * wei[p];
        }
    }
    }
    if (!clip) return;
    REAL *f = (REAL *)malloc(sizeof(REAL) * (n > 0 ? n : 1));
    for (long p = 0; p < n; p++) f[p] = res[p] < 0 ? -res[p] : res[p];
    qsort(f, n, sizeof(REAL), FN(orc_cmp));
    long pos = (long)(int)((97L * nt * ntr) / 100);
    REAL pclip = f[pos];
    free(f);
    for (long p = 0; p < n; p++) {
        REAL a = res[p] < 0 ? -res[p] : res[p];
        if (a >= pclip) res[p] = 0.0;
    }
}

/* ---- Elastic 2D misfits and residuals (FwiElastic2D::computeMisfit /
 * computeResiduals, fwi.cpp:2852-3773) over the P, Ux and Uz records of one
 * shot, [ntr][nt] each; data weights optional per component (NULL = not
 * set); `valid[i]` nonzero when trace i's Ux map is inside the model. */

/* Joint shaping filter of ADAPTIVE_SINGLE: the Ux and Uz traces (weighted)
 * back to back as one 2 nt sequence, auto- and crosscorrelation over nt lags,
 * whitening, Levinson.  w: 2nt modwei, 2nt recwei, 2nt reversed copy, nt
 * autocorr, nt crosscorr, nt spiker; returns the squared norm (1 when 0). */
static REAL FN(single_shaper_)(const REAL *mx, const REAL *rx, const REAL *wx, const REAL *mz, const REAL *rz,
                               const REAL *wz, int nt, REAL *w, REAL *shaper)
{
    REAL *modwei = w, *recwei = w + 2 * nt, *xr = w + 4 * nt, *autoc = w + 6 * nt, *cross = w + 7 * nt,
         *spiker = w + 8 * nt;
    for (int it = 0; it < nt; it++) {
        modwei[it] = wx ? mx[it] * wx[it] : mx[it];
        recwei[it] = wx ? rx[it] * wx[it] : rx[it];
        modwei[nt + it] = wz ? mz[it] * wz[it] : mz[it];
        recwei[nt + it] = wz ? rz[it] * wz[it] : rz[it];
    }
    FN(xcor_)(2 * nt, 0, recwei, 2 * nt, 0, recwei, nt, 0, autoc, xr);
    FN(xcor_)(2 * nt, 0, recwei, 2 * nt, 0, modwei, nt, 0, cross, xr);
    if (autoc[0] == 0.0) autoc[0] = 1.0;
    autoc[0] = (REAL)((double)autoc[0] * (1.0 + 0.01));
    for (int it = 0; it < nt; it++) shaper[it] = 0.0;
    FN(stoep_)(nt, autoc, cross, shaper, spiker);
    REAL norm = 0.0;
    for (int it = 0; it < nt; it++) norm += shaper[it] * shaper[it];
    if (norm == 0.0) norm = 1.0;
    return norm;
}

static int FN(finite_)(REAL v) { return !isnan(v) && !isinf(v); }

/* Misfit: type 0 DIFFERENCE 0.5*(rp^2+rx^2+rz^2) (weighted differences),
 * 1 CORRELATION sum of -mod*rec/(n1 n2) (weighted, non-finite terms 0) over
 * the three records, 2/3 ADAPTIVE_GAUSS/LINEAR -/+ 0.5*(Hx sx)^2/|sx|^2 +
 * (Hz sz)^2/|sz|^2 over Ux and Uz, 4 ADAPTIVE_SINGLE with one joint shaper;
 * accumulated in T with the reference's promotions. */
REAL FN(orc_misfit_e)(int type, const REAL *modp, const REAL *recp, const REAL *modx, const REAL *recx,
                      const REAL *modz, const REAL *recz, const REAL *weip, const REAL *weix, const REAL *weiz,
                      const int *valid, int ntr, int nt)
{
    REAL misfit = 0.0;
    REAL *w = (REAL *)malloc(sizeof(REAL) * 12 * (nt > 0 ? nt : 1));
    REAL *shx = w + 10 * nt, *shz = w + 11 * nt;
    for (int i = 0; i < ntr; i++) {
        if (!valid[i]) continue;
        const long o = (long)i * nt;
        const REAL *mp = modp + o, *rp = recp + o, *mx = modx + o, *rx = recx + o, *mz = modz + o, *rz = recz + o;
        const REAL *wp = weip ? weip + o : NULL, *wx = weix ? weix + o : NULL, *wz = weiz ? weiz + o : NULL;
        if (type == 0) {
            for (int it = 0; it < nt; it++) {
                REAL a = mp[it] - rp[it], b = mx[it] - rx[it], c = mz[it] - rz[it];
                if (wp) a *= wp[it];
                if (wx) b *= wx[it];
                if (wz) c *= wz[it];
                misfit = (REAL)((double)misfit + 0.5 * (double)(a * a + b * b + c * c));
            }
        } else if (type == 1) {
            REAL p1 = 0.0, p2 = 0.0, x1 = 0.0, x2 = 0.0, z1 = 0.0, z2 = 0.0;
            for (int it = 0; it < nt; it++) {
                p1 += mp[it] * mp[it];
                p2 += rp[it] * rp[it];
                x1 += mx[it] * mx[it];
                x2 += rx[it] * rx[it];
                z1 += mz[it] * mz[it];
                z2 += rz[it] * rz[it];
            }
            p1 = (REAL)sqrt((double)p1);
            p2 = (REAL)sqrt((double)p2);
            if (p1 == 0) p1 = 1.0;
            if (p2 == 0) p2 = 1.0;
            x1 = (REAL)sqrt((double)x1);
            x2 = (REAL)sqrt((double)x2);
            if (x1 == 0) x1 = 1.0;
            if (x2 == 0) x2 = 1.0;
            z1 = (REAL)sqrt((double)z1);
            z2 = (REAL)sqrt((double)z2);
            if (z1 == 0) z1 = 1.0;
            if (z2 == 0) z2 = 1.0;
            for (int it = 0; it < nt; it++) {
                REAL a = (REAL)((-1.0) * (double)(mp[it] * rp[it] / (p1 * p2)));
                REAL b = (REAL)((-1.0) * (double)(mx[it] * rx[it] / (x1 * x2)));
                REAL c = (REAL)((-1.0) * (double)(mz[it] * rz[it] / (z1 * z2)));
                if (wp) a *= wp[it];
                if (wx) b *= wx[it];
                if (wz) c *= wz[it];
                if (!FN(finite_)(a)) a = 0.0;
                if (!FN(finite_)(b)) b = 0.0;
                if (!FN(finite_)(c)) c = 0.0;
                misfit += (a + b + c);
            }
        } else if (type == 2 || type == 3) {
            REAL nx = FN(adaptive_shaper_)(mx, rx, wx, nt, w, shx);
            REAL nz = FN(adaptive_shaper_)(mz, rz, wz, nt, w, shz);
            for (int it = 0; it < nt; it++) {
                REAL h = FN(adaptive_h_)(type, it, nt);
                REAL a = h * shx[it], b = h * shz[it];
                REAL v = a * a / nx + b * b / nz;
                misfit = (REAL)((double)misfit + (type == 2 ? -0.5 : 0.5) * (double)v);
            }
        } else {
            REAL nx = FN(single_shaper_)(mx, rx, wx, mz, rz, wz, nt, w, shx);
            for (int it = 0; it < nt; it++) {
                REAL a = FN(adaptive_h_)(3, it, nt) * shx[it];
                misfit = (REAL)((double)misfit + 0.5 * (double)(a * a / nx));
            }
        }
    }
    free(w);
    return misfit;
}

/* Value at the 97th percentile of |a| (element (int)(97*nt*ntr/100) of the
 * sorted magnitudes, fwi.cpp:3380-3390). */
static REAL FN(pclip_)(const REAL *a, int ntr, int nt)
{
    const long n = (long)ntr * nt;
    REAL *f = (REAL *)malloc(sizeof(REAL) * (n > 0 ? n : 1));
    for (long p = 0; p < n; p++) f[p] = a[p] < 0 ? -a[p] : a[p];
    qsort(f, n, sizeof(REAL), FN(orc_cmp));
    REAL v = f[(long)(int)((97L * nt * ntr) / 100)];
    free(f);
    return v;
}

/* Combined clip of Ux and Uz: the larger of the two percentiles; both
 * components zeroed where |resx| reaches it (|resz| is not tested), then
 * with `nanfix` non-finite values zeroed (fwi.cpp:3372-3406). */
static void FN(clip_xz_)(REAL *rx, REAL *rz, int ntr, int nt, int nanfix)
{
    const long n = (long)ntr * nt;
    if (n == 0) return;
    REAL pc = FN(pclip_)(rx, ntr, nt), pz = FN(pclip_)(rz, ntr, nt);
    if (pz > pc) pc = pz;
    for (long p = 0; p < n; p++) {
        REAL a = rx[p] < 0 ? -rx[p] : rx[p];
        if (a >= pc) {
            rx[p] = 0.0;
            rz[p] = 0.0;
        }
        if (nanfix) {
            if (!FN(finite_)(rx[p])) rx[p] = 0.0;
            if (!FN(finite_)(rz[p])) rz[p] = 0.0;
        }
    }
}

/* Residuals of every trace: type 0 (mod-rec)*w^2; 1 CORRELATION
 * -(rec/(n1 n2) - mod/(n1 n1) * sum(mod rec)/(n1 n2)) * w^2, combined Ux/Uz
 * clip with non-finite values zeroed, P clipped on its own; 2/3 the
 * ADAPTIVE residual of Ux and Uz separately, combined clip; 4 the joint
 * ADAPTIVE_SINGLE residual, no clip.  Types 2-4 leave P at zero. */
void FN(orc_residual_e)(int type, const REAL *modp, const REAL *recp, const REAL *modx, const REAL *recx,
                        const REAL *modz, const REAL *recz, const REAL *weip, const REAL *weix, const REAL *weiz,
                        REAL *resp, REAL *resx, REAL *resz, int ntr, int nt)
{
    const long n = (long)ntr * nt;
    const REAL *mod[3] = {modp, modx, modz}, *rec[3] = {recp, recx, recz}, *wei[3] = {weip, weix, weiz};
    REAL *res[3] = {resp, resx, resz};
    if (type == 0) {
        for (int c = 0; c < 3; c++)
            for (long p = 0; p < n; p++) {
                res[c][p] = mod[c][p] - rec[c][p];
                if (wei[c]) res[c][p] *= wei[c][p] * wei[c][p];
            }
        return;
    }
    if (type == 1) {
        for (int c = 0; c < 3; c++)
            for (int i = 0; i < ntr; i++) {
                const REAL *m = mod[c] + (long)i * nt, *r = rec[c] + (long)i * nt;
                REAL n1 = 0.0, n2 = 0.0, n3 = 0.0;
                for (int it = 0; it < nt; it++) {
                    n1 += m[it] * m[it];
                    n2 += r[it] * r[it];
                    n3 += m[it] * r[it];
                }
                n1 = (REAL)sqrt((double)n1);
                n2 = (REAL)sqrt((double)n2);
                if (n1 == 0) n1 = 1.0;
                if (n2 == 0) n2 = 1.0;
                n3 /= (n1 * n2);
                for (int it = 0; it < nt; it++) {
                    long p = (long)i * nt + it;
                    res[c][p] = (REAL)((-1.0) * (double)((r[it] / (n1 * n2)) - (m[it] / (n1 * n1)) * n3));
                    if (wei[c]) res[c][p] *= wei[c][p] * wei[c][p];
                }
            }
        FN(clip_xz_)(resx, resz, ntr, nt, 1);
        if (n > 0) {
            REAL pc = FN(pclip_)(resp, ntr, nt);
            for (long p = 0; p < n; p++) {
                REAL a = resp[p] < 0 ? -resp[p] : resp[p];
                if (a >= pc) resp[p] = 0.0;
                if (!FN(finite_)(resp[p])) resp[p] = 0.0;
            }
        }
        return;
    }
    for (long p = 0; p < n; p++) resp[p] = 0.0;
    REAL *w = (REAL *)malloc(sizeof(REAL) * 11 * (nt > 0 ? nt : 1)), *shaper = w + 10 * nt;
    if (type == 2 || type == 3) {
        for (int c = 1; c < 3; c++)
            for (int i = 0; i < ntr; i++)
                FN(adaptive_res_)(type, mod[c] + (long)i * nt, rec[c] + (long)i * nt,
                                  wei[c] ? wei[c] + (long)i * nt : NULL, res[c] + (long)i * nt, nt, w, shaper, 1);
        free(w);
        FN(clip_xz_)(resx, resz, ntr, nt, 0);
        return;
    }
    for (int i = 0; i < ntr; i++) {
        const long o = (long)i * nt;
        const REAL *wx = weix ? weix + o : NULL, *wz = weiz ? weiz + o : NULL;
        REAL norm = FN(single_shaper_)(modx + o, recx + o, wx, modz + o, recz + o, wz, nt, w, shaper);
        REAL misfit = 0.0;
        for (int it = 0; it < nt; it++) {
            REAL h = FN(adaptive_h_)(3, it, nt);
            misfit = (REAL)((double)misfit + 0.5 * (double)(h * shaper[it]) * (double)(h * shaper[it]) / (double)norm);
        }
        REAL *cross = w + 7 * nt;
        for (int it = 0; it < nt; it++) {
            REAL h = FN(adaptive_h_)(3, it, nt);
            cross[it] = (REAL)(((double)(h * h) - 2.0 * (double)misfit) * (double)shaper[it] / (double)norm);
        }
        FN(stoep_)(nt, w + 6 * nt, cross, shaper, w + 8 * nt);
        /* the joint residual (2 nt) overwrites the weighted model traces */
        FN(conv_)(nt, 0, shaper, 2 * nt, 0, w + 2 * nt, 2 * nt, 0, w);
        for (int it = 0; it < nt; it++) {
            resx[o + it] = wx ? w[it] * wx[it] : w[it];
            resz[o + it] = wz ? w[nt + it] * wz[it] : w[nt + it];
        }
    }
    free(w);
}

/* DIFFERENCE misfit 0.5*sum(r^2) over traces with valid maps, accumulated
 * in T (fwi.cpp:574-608); `valid[i]` nonzero when map x,z >= 0. */
REAL FN(orc_misfit_diff)(const REAL *mod, const REAL *rec, const int *valid, int ntr, int nt)
{
    REAL misfit = 0.0;
    for (int i = 0; i < ntr; i++) {
        if (!valid[i]) continue;
        for (int it = 0; it < nt; it++) {
            REAL res = mod[(long)i * nt + it] - rec[(long)i * nt + it];
            misfit = (REAL)((double)misfit + 0.5 * (double)res * (double)res);
        }
    }
    return misfit;
}

/* DIFFERENCE residual scaling*(mod-rec), scaling=-1 for VX/VZ fields
 * (fwi.cpp:772-800). */
void FN(orc_residual_diff)(const REAL *mod, const REAL *rec, REAL *res, int ntr, int nt, int scaling)
{
    long n = (long)ntr * nt;
    for (long p = 0; p < n; p++) res[p] = scaling * (mod[p] - rec[p]);
}

/* ---------------------------------------------------------------------- */
/* FwiElastic3D (fwi.cpp:5046-5291, 5553-6196)                              */
/* ---------------------------------------------------------------------- */

/* crossCorr of run_optimal (fwi.cpp:5046-5291), no domain decomposition:
 * source fields ws* / fs* and receiver fields wr* / rs* on the padded grid
 * (pad lpml), Rx/Ry/Rz padded; lambda (vpgrad), mu (vsgrad) and rho kernels
 * over cells 1..n-2 of the unpadded grid, with the reference's operand order
 * and T / double promotions.  Kept as the reference has them: all four shear
 * terms of each pair read the receiver stress at the cell itself, and the
 * second, fourth and sixth z terms of the rho kernel read wrz one cell
 * back in y with Rz one cell back in z. */
void FN(orc_e3d_crosscorr)(const REAL *wsx, const REAL *wsy, const REAL *wsz, const REAL *fsxx, const REAL *fsyy,
                           const REAL *fszz, const REAL *fsyz, const REAL *fsxz, const REAL *fsxy, const REAL *wrx,
                           const REAL *wry, const REAL *wrz, const REAL *rsxx, const REAL *rsyy, const REAL *rszz,
                           const REAL *rsyz, const REAL *rsxz, const REAL *rsxy, const REAL *Rx, const REAL *Ry,
                           const REAL *Rz, int nx, int ny, int nz, int lpml, REAL dx, REAL dy, REAL dz, REAL *lam,
                           REAL *mu, REAL *rho)
{
    const long nxp = nx + 2 * lpml, nxyp = nxp * (ny + 2 * lpml);
#define K(x, y, z) ((long)(z) * nxyp + (long)(y) * nxp + (x))
#define SUBD(v, t) v = (REAL)((double)v - 0.25 * (double)(t) * (double)rs_c)
#define ADDU(u, w, r, d, h) u = (REAL)((double)u + 0.5 * (double)(w) * (double)(r) * (double)(d) / (double)(h))
    for (int iz = 1; iz < nz - 1; iz++)
        for (int iy = 1; iy < ny - 1; iy++)
            for (int ix = 1; ix < nx - 1; ix++) {
                const int x = ix + lpml, y = iy + lpml, z = iz + lpml;
                const long c = K(x, y, z), m = ((long)iz * ny + iy) * nx + ix;
                REAL msxx = (wsx[c] - wsx[K(x - 1, y, z)]) / dx;
                REAL msyy = (wsy[c] - wsy[K(x, y - 1, z)]) / dy;
                REAL mszz = (wsz[c] - wsz[K(x, y, z - 1)]) / dz;
                lam[m] -= (msxx + msyy + mszz) * (rsxx[c] + rsyy[c] + rszz[c]);
                REAL v = mu[m], t, rs_c;
                rs_c = rsyz[c];
                t = (wsy[K(x, y - 1, z)] - wsy[K(x, y - 1, z - 1)]) / dz;
                t += (wsz[K(x, y, z - 1)] - wsz[K(x, y - 1, z - 1)]) / dy;
                SUBD(v, t);
                t = (wsy[c] - wsy[K(x, y, z - 1)]) / dz;
                t += (wsz[K(x, y + 1, z - 1)] - wsz[K(x, y, z - 1)]) / dy;
                SUBD(v, t);
                t = (wsy[K(x, y - 1, z + 1)] - wsy[K(x, y - 1, z)]) / dz;
                t += (wsz[c] - wsz[K(x, y - 1, z)]) / dy;
                SUBD(v, t);
                t = (wsy[K(x, y, z + 1)] - wsy[c]) / dz;
                t += (wsz[K(x, y + 1, z)] - wsz[c]) / dy;
                SUBD(v, t);
                rs_c = rsxz[c];
                t = (wsx[K(x - 1, y, z)] - wsx[K(x - 1, y, z - 1)]) / dz;
                t += (wsz[K(x, y, z - 1)] - wsz[K(x - 1, y, z - 1)]) / dx;
                SUBD(v, t);
                t = (wsx[c] - wsx[K(x, y, z - 1)]) / dz;
                t += (wsz[K(x + 1, y, z - 1)] - wsz[K(x, y, z - 1)]) / dx;
                SUBD(v, t);
                t = (wsx[K(x - 1, y, z + 1)] - wsx[K(x - 1, y, z)]) / dz;
                t += (wsz[c] - wsz[K(x - 1, y, z)]) / dx;
                SUBD(v, t);
                t = (wsx[K(x, y, z + 1)] - wsx[c]) / dz;
                t += (wsz[K(x + 1, y, z)] - wsz[c]) / dx;
                SUBD(v, t);
                rs_c = rsxy[c];
                t = (wsy[K(x, y - 1, z)] - wsy[K(x - 1, y - 1, z)]) / dx;
                t += (wsx[K(x - 1, y, z)] - wsx[K(x - 1, y - 1, z)]) / dy;
                SUBD(v, t);
                t = (wsy[K(x + 1, y - 1, z)] - wsy[K(x, y - 1, z)]) / dx;
                t += (wsx[c] - wsx[K(x, y - 1, z)]) / dy;
                SUBD(v, t);
                t = (wsy[c] - wsy[K(x - 1, y, z)]) / dx;
                t += (wsx[K(x - 1, y + 1, z)] - wsx[K(x - 1, y, z)]) / dy;
                SUBD(v, t);
                t = (wsy[K(x + 1, y, z)] - wsy[c]) / dx;
                t += (wsx[K(x, y + 1, z)] - wsx[c]) / dy;
                SUBD(v, t);
                mu[m] = (REAL)((double)v - (2.0 * (double)msxx * (double)rsxx[c] + 2.0 * (double)msyy * (double)rsyy[c] +
                                            2.0 * (double)mszz * (double)rszz[c]));
                const long xm = K(x - 1, y, z), ym = K(x, y - 1, z), zm = K(x, y, z - 1);
                REAL ux = (REAL)(0.5 * (double)wrx[c] * (double)Rx[c] * (double)(fsxx[K(x + 1, y, z)] - fsxx[c]) /
                                 (double)dx);
                ADDU(ux, wrx[xm], Rx[xm], fsxx[c] - fsxx[xm], dx);
                ADDU(ux, wrx[c], Rx[c], fsxy[c] - fsxy[ym], dy);
                ADDU(ux, wrx[xm], Rx[xm], fsxy[xm] - fsxy[K(x - 1, y - 1, z)], dy);
                ADDU(ux, wrx[c], Rx[c], fsxz[c] - fsxz[zm], dz);
                ADDU(ux, wrx[xm], Rx[xm], fsxz[xm] - fsxz[K(x - 1, y, z - 1)], dz);
                REAL uy = (REAL)(0.5 * (double)wry[c] * (double)Ry[c] * (double)(fsxy[c] - fsxy[xm]) / (double)dx);
                ADDU(uy, wry[ym], Ry[ym], fsxy[ym] - fsxy[K(x - 1, y - 1, z)], dx);
                ADDU(uy, wry[c], Ry[c], fsyy[K(x, y + 1, z)] - fsyy[c], dy);
                ADDU(uy, wry[ym], Ry[ym], fsyy[c] - fsyy[ym], dy);
                ADDU(uy, wry[c], Ry[c], fsyz[c] - fsyz[zm], dz);
                ADDU(uy, wry[ym], Ry[ym], fsyz[ym] - fsyz[K(x, y - 1, z - 1)], dz);
                REAL uz = (REAL)(0.5 * (double)wrz[c] * (double)Rz[c] * (double)(fsxz[c] - fsxz[xm]) / (double)dx);
                ADDU(uz, wrz[ym], Rz[zm], fsxz[zm] - fsxz[K(x - 1, y, z - 1)], dx);
                ADDU(uz, wrz[c], Rz[c], fsyz[c] - fsyz[ym], dy);
                ADDU(uz, wrz[ym], Rz[zm], fsyz[zm] - fsyz[K(x, y - 1, z - 1)], dy);
                ADDU(uz, wrz[c], Rz[c], fszz[K(x, y, z + 1)] - fszz[c], dz);
                ADDU(uz, wrz[ym], Rz[zm], fszz[c] - fszz[zm], dz);
                rho[m] -= (ux + uy + uz);
            }
#undef ADDU
#undef SUBD
#undef K
}

/* FwiElastic3D::scaleGrad (fwi.cpp:5553-5630): lambda/mu/rho kernels ->
 * Vp/Vs/rho gradients in place.  The rho gradient is vp^2 lambda + vs^2 mu
 * + rho as the reference forms it. */
void FN(orc_e3d_scalegrad)(const REAL *Vp, const REAL *Vs, const REAL *Rho, long n, REAL *vpgrad, REAL *vsgrad,
                           REAL *rhograd)
{
    for (long k = 0; k < n; k++) {
        REAL vpscale = (REAL)(2.0 * (double)Rho[k] * (double)Vp[k]);
        REAL vsscale = (REAL)(2.0 * (double)Rho[k] * (double)Vs[k]);
        REAL rhoscale1 = Vp[k] * Vp[k], rhoscale2 = Vs[k] * Vs[k];
        REAL lambda = vpgrad[k], mu = vsgrad[k], rho = rhograd[k];
        vpgrad[k] = vpscale * lambda;
        vsgrad[k] = (REAL)((double)(vsscale * mu) - 2.0 * (double)vsscale * (double)lambda);
        rhograd[k] = rhoscale1 * lambda + rhoscale2 * mu + rho;
    }
}

/* Shapers of the three components of one trace for ADAPTIVE_GAUSS
 * (fwi.cpp:5773-5830, 6047-6100), with the reference's weighted inputs:
 * weighted x leaves the record input zero, weighted y the model input zero
 * (both are only ever assigned the other one), z takes both weighted.
 * w: 3 * 8 nt (per component 6 nt of adaptive_shaper_, the shaper, and a
 * weighted-input copy); returns the squared norms in nrm. */
static void FN(e3_shapers_)(const REAL *const *m, const REAL *const *r, const REAL *const *wt, int nt, REAL *w,
                            REAL *nrm)
{
    for (int c = 0; c < 3; c++) {
        REAL *wc = w + (long)c * 8 * nt, *sh = wc + 6 * nt, *in = wc + 7 * nt;
        if (wt[c] && c < 2) {
            /* one input weighted, the other zero */
            for (int it = 0; it < nt; it++) in[it] = c == 0 ? m[c][it] * wt[c][it] : r[c][it] * wt[c][it];
            REAL *zero = (REAL *)calloc(nt > 0 ? nt : 1, sizeof(REAL));
            nrm[c] = c == 0 ? FN(adaptive_shaper_)(in, zero, NULL, nt, wc, sh)
                            : FN(adaptive_shaper_)(zero, in, NULL, nt, wc, sh);
            free(zero);
        } else {
            nrm[c] = FN(adaptive_shaper_)(m[c], r[c], wt[c], nt, wc, sh);
        }
    }
}

/* computeMisfit (fwi.cpp:5632-5899) of the Ux, Uy, Uz records, every trace:
 * DIFFERENCE and the default branch 0.5 (rx^2 + ry^2 + rz^2) of the weighted
 * differences, CORRELATION -sum mod rec / (|mod||rec|) weighted (added in
 * T), ADAPTIVE_GAUSS 0.5 sum (H s)^2 / |s|^2 over the three shapers. */
REAL FN(orc_misfit_e3)(int type, const REAL *modx, const REAL *recx, const REAL *mody, const REAL *recy,
                       const REAL *modz, const REAL *recz, const REAL *weix, const REAL *weiy, const REAL *weiz, int ntr,
                       int nt)
{
    REAL misfit = 0.0;
    REAL *w = (REAL *)malloc(sizeof(REAL) * 24 * (nt > 0 ? nt : 1));
    for (int i = 0; i < ntr; i++) {
        const long o = (long)i * nt;
        const REAL *m[3] = {modx + o, mody + o, modz + o}, *r[3] = {recx + o, recy + o, recz + o};
        const REAL *wt[3] = {weix ? weix + o : NULL, weiy ? weiy + o : NULL, weiz ? weiz + o : NULL};
        if (type == 1) {
            REAL n1[3], n2[3];
            for (int c = 0; c < 3; c++) {
                REAL a = 0.0, b = 0.0;
                for (int it = 0; it < nt; it++) {
                    a += m[c][it] * m[c][it];
                    b += r[c][it] * r[c][it];
                }
                n1[c] = (REAL)sqrt((double)a);
                n2[c] = (REAL)sqrt((double)b);
                if (n1[c] == 0) n1[c] = 1.0;
                if (n2[c] == 0) n2[c] = 1.0;
            }
            for (int it = 0; it < nt; it++) {
                REAL v[3];
                for (int c = 0; c < 3; c++) {
                    v[c] = (REAL)((-1.0) * (double)(m[c][it] * r[c][it] / (n1[c] * n2[c])));
                    if (wt[c]) v[c] *= wt[c][it];
                }
                misfit += (v[0] + v[1] + v[2]);
            }
        } else if (type == 2) {
            REAL nrm[3];
            FN(e3_shapers_)(m, r, wt, nt, w, nrm);
            const REAL *sx = w + 6 * nt, *sy = w + 14 * nt, *sz = w + 22 * nt;
            for (int it = 0; it < nt; it++) {
                REAL h = FN(adaptive_h_)(2, it, nt);
                REAL a = h * sx[it], b = h * sy[it], c = h * sz[it];
                misfit = (REAL)((double)misfit + 0.5 * (double)(a * a / nrm[0] + b * b / nrm[1] + c * c / nrm[2]));
            }
        } else {
            for (int it = 0; it < nt; it++) {
                REAL v[3];
                for (int c = 0; c < 3; c++) {
                    v[c] = m[c][it] - r[c][it];
                    if (wt[c]) v[c] *= wt[c][it];
                }
                misfit = (REAL)((double)misfit + 0.5 * (double)(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
            }
        }
    }
    free(w);
    return misfit;
}

/* computeResiduals (fwi.cpp:5901-6196): DIFFERENCE (mod - rec) w^2 from
 * sample 0; CORRELATION -(rec/(n1 n2) - mod/(n1 n1) n3) w^2 and the
 * default branch (mod - rec) w^2 from sample 1 (sample 0 left as it is);
 * ADAPTIVE_GAUSS per component -(H^2 - 2 misfit) s / |s|^2, solved with the
 * record's autocorrelation, convolved with the unweighted record, not
 * weighted. */
void FN(orc_residual_e3)(int type, const REAL *modx, const REAL *recx, const REAL *mody, const REAL *recy,
                         const REAL *modz, const REAL *recz, const REAL *weix, const REAL *weiy, const REAL *weiz,
                         REAL *resx, REAL *resy, REAL *resz, int ntr, int nt)
{
    REAL *w = (REAL *)malloc(sizeof(REAL) * 24 * (nt > 0 ? nt : 1));
    for (int i = 0; i < ntr; i++) {
        const long o = (long)i * nt;
        const REAL *m[3] = {modx + o, mody + o, modz + o}, *r[3] = {recx + o, recy + o, recz + o};
        const REAL *wt[3] = {weix ? weix + o : NULL, weiy ? weiy + o : NULL, weiz ? weiz + o : NULL};
        REAL *rs[3] = {resx + o, resy + o, resz + o};
        if (type == 2) {
            REAL nrm[3];
            FN(e3_shapers_)(m, r, wt, nt, w, nrm);
            for (int c = 0; c < 3; c++) {
                REAL *wc = w + (long)c * 8 * nt, *sh = wc + 6 * nt, mis = 0.0;
                for (int it = 0; it < nt; it++) {
                    REAL h = FN(adaptive_h_)(2, it, nt);
                    mis = (REAL)((double)mis + 0.5 * (double)(h * sh[it]) * (double)(h * sh[it]) / (double)nrm[c]);
                }
                for (int it = 0; it < nt; it++) {
                    REAL h = FN(adaptive_h_)(2, it, nt);
                    rs[c][it] = (REAL)(-1.0 * ((double)(h * h) - 2.0 * (double)mis) * (double)sh[it] / (double)nrm[c]);
                }
                FN(stoep_)(nt, wc + 2 * nt, rs[c], sh, wc + 4 * nt);
                FN(conv_)(nt, 0, sh, nt, 0, r[c], nt, 0, rs[c]);
                /* no weighting: the reference's weight multiply (fwi.cpp:6143-6154)
                 * sits after the sample loop and hits index nt, the next trace's
                 * sample 0, which that trace then rewrites */
            }
            continue;
        }
        REAL n1[3] = {1, 1, 1}, n3[3] = {0, 0, 0}, n2[3] = {1, 1, 1};
        if (type == 1)
            for (int c = 0; c < 3; c++) {
                REAL a = 0.0, b = 0.0, x = 0.0;
                for (int it = 0; it < nt; it++) {
                    a += m[c][it] * m[c][it];
                    b += r[c][it] * r[c][it];
                    x += m[c][it] * r[c][it];
                }
                n1[c] = (REAL)sqrt((double)a);
                n2[c] = (REAL)sqrt((double)b);
                if (n1[c] == 0) n1[c] = 1.0;
                if (n2[c] == 0) n2[c] = 1.0;
                n3[c] = x / (n1[c] * n2[c]);
            }
        for (int it = type == 0 ? 0 : 1; it < nt; it++)
            for (int c = 0; c < 3; c++) {
                REAL v = type == 1 ? (REAL)((-1.0) * (double)((r[c][it] / (n1[c] * n2[c])) - (m[c][it] / (n1[c] * n1[c])) * n3[c]))
                                   : m[c][it] - r[c][it];
                if (wt[c]) v *= wt[c][it] * wt[c][it];
                rs[c][it] = v;
            }
    }
    free(w);
}

#undef FN
#undef CAT
#undef CAT_
