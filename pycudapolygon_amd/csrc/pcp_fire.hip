// FIRE minimization on the device (Model::minimizeFIREStep / minimizeFIRE,
// src/model.cpp:405-504).  The FIRE kernels of FIRE.h are absent from the reference
// (src/model.cu:15); they are written here from the launch contract model.cu:576-608
// and Bitzek et al., PRL 97, 170201 (2006):
//   updatePositionAndVelocity:  vh = v + (0.5*dt)*F;  x = frac(x + dt*vh);  v = vh
//   updateVelocity:             v = v + (0.5*dt)*F
//   bendVelocityTowardsForce:   P = F.v, |v|, |F|;  v = (1-a)*v + ((a*|v|)/|F|)*F
//   rederiveVelocity:           v = v + wrap(x_shaken - x_unshaken)/dt
// All per-packing state (dt, alpha, nPos, energies, flags) lives on the device; the
// kernels skip packings that have converged, so P independent packings advance in the
// same launches.  The host reads one small status block per step.
//
// This file is compiled as part of pcp_kernels.hip (one translation unit).

namespace pcp {

enum { FL_FAIL = 0, FL_ROLLBACK = 1, FL_ACCEPT = 2, FL_ZEROV = 3, FL_LASTSHAKE = 4, FL_ACTIVEBAK = 5,
       FL_NF = 8 };
enum { FS_SCALE = 0, FS_ONEMA = 1, FS_MAXF = 2, FS_DTB = 3, FS_ALB = 4, FS_NPB = 5, FS_STEPB = 6,
       FS_MAXFB = 7, FS_NF = 16 };

__global__ void k_fire_init(int P, double dtv, double alpha0, double* dt, double* alpha, int* nPos,
                            int* steps, int* active, int* flags, double* fs) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    dt[p] = dtv;
    alpha[p] = alpha0;
    nPos[p] = 0;
    steps[p] = 0;
    active[p] = 1;
    for (int k = 0; k < FL_NF; ++k) flags[FL_NF * p + k] = 0;
    fs[FS_NF * p + FS_MAXF] = 0.0;
}

// Step start.  Per packing (thread p < P): backups for rollback / retry, flag reset,
// E_pre (model.cpp:411-412).  Per vertex (thread i < V): velocity-Verlet half kick and
// drift, with backups of x and v.  XPBD: a full kick with the force before the springs
// (Fe; the springs are integrated by the XPBD projection that follows), backed up in febak.
template <bool XPBD>
__global__ void k_fire_begin(int V, int P, const int* __restrict__ shapeId, const int* __restrict__ polyPack,
                             const int* __restrict__ active, const double* __restrict__ dt,
                             const double* __restrict__ alpha, const int* __restrict__ nPos,
                             const int* __restrict__ steps, const double* __restrict__ energy,
                             double* __restrict__ epre, int* __restrict__ flags, double* __restrict__ fs,
                             int* __restrict__ shakeIt, int* __restrict__ summary, double2* __restrict__ x,
                             double2* __restrict__ v, const double2* __restrict__ F, double2* __restrict__ xbak,
                             double2* __restrict__ vbak, double2* __restrict__ xpre, double2* __restrict__ fbak,
                             const double2* __restrict__ Fe, double2* __restrict__ febak) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 4) summary[i] = 0;
    if (i < P) {
        int p = i;
        int* fl = flags + FL_NF * p;
        fl[FL_FAIL] = 0;
        fl[FL_ROLLBACK] = 0;
        fl[FL_ACCEPT] = 0;
        fl[FL_ZEROV] = 0;
        fl[FL_ACTIVEBAK] = active[p];
        shakeIt[p] = 0;
        shakeIt[P + p] = 0;
        if (active[p]) {
            epre[p] = energy[p];
            fs[FS_NF * p + FS_DTB] = dt[p];
            fs[FS_NF * p + FS_ALB] = alpha[p];
            fs[FS_NF * p + FS_NPB] = (double)nPos[p];
            fs[FS_NF * p + FS_STEPB] = (double)steps[p];
            fs[FS_NF * p + FS_MAXFB] = fs[FS_NF * p + FS_MAXF];
        }
    }
    if (i >= V) return;
    int p = polyPack[shapeId[i]];
    if (!active[p]) return;
    double h = dt[p];
    double2 xi = x[i], vi = v[i], fi = F[i];
    xbak[i] = xi;
    vbak[i] = vi;
    if (fbak) fbak[i] = fi;
    double vhx, vhy;
    if (XPBD) {
        double2 fe = Fe[i];
        if (febak) febak[i] = fe;
        vhx = vi.x + h * fe.x;
        vhy = vi.y + h * fe.y;
    } else {
        vhx = vi.x + (0.5 * h) * fi.x;
        vhy = vi.y + (0.5 * h) * fi.y;
    }
    double px = xi.x + h * vhx, py = xi.y + h * vhy;
    px = px - floor(px);
    py = py - floor(py);
    x[i] = make_double2(px, py);
    xpre[i] = make_double2(px, py);
    v[i] = make_double2(vhx, vhy);
}

// after SHAKE: restore on failure (model.cpp:431-435) else rederive (model.cpp:437)
__global__ void k_fire_post_shake(int V, int P, int shakeIter, const int* __restrict__ shapeId,
                                  const int* __restrict__ polyPack, const int* __restrict__ packVert,
                                  const int* __restrict__ active, const double* __restrict__ dt,
                                  const int* __restrict__ shakeIt, double2* __restrict__ x,
                                  double2* __restrict__ v, const double2* __restrict__ xbak,
                                  const double2* __restrict__ xpre, int* __restrict__ flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    int p = polyPack[shapeId[i]];
    if (!active[p]) return;
    int it1 = shakeIt[p], it2 = shakeIt[P + p];
    bool retried = it1 >= shakeIter;
    bool fail = retried && it2 >= shakeIter * 10;
    if (fail) {
        x[i] = xbak[i];
        v[i] = make_double2(0.0, 0.0);
    } else {
        double h = dt[p];
        double2 xi = x[i], xp = xpre[i], vi = v[i];
        v[i] = make_double2(vi.x + wrap(xi.x - xp.x) / h, vi.y + wrap(xi.y - xp.y) / h);
    }
    if (i == packVert[p]) {
        flags[FL_NF * p + FL_FAIL] = fail ? 1 : 0;
        flags[FL_NF * p + FL_LASTSHAKE] = retried ? it2 : it1;
    }
}

// The FIRE scalars of one packing held in registers by the thread that updates them: read
// together at a kernel's start (one memory latency, overlapping the kernel's reductions) and
// written back once, so a packing's bookkeeping is not a chain of dependent loads and stores.
struct FireRegs {
    int act, nPos, steps;
    int fl[FL_NF];
    double dt, alpha, energy, epre;
    double fsScale, fsOneMa, fsMaxF, fsMaxFB;
};
__device__ __forceinline__ FireRegs fire_load(int p, const int* active, const int* flags, const double* dt,
                                              const double* alpha, const int* nPos, const int* steps,
                                              const double* energy, const double* epre, const double* fs) {
    FireRegs S;
    S.act = active[p];
#pragma unroll
    for (int k = 0; k < FL_NF; ++k) S.fl[k] = flags[FL_NF * p + k];
    S.dt = dt[p];
    S.alpha = alpha[p];
    S.nPos = nPos[p];
    S.steps = steps[p];
    S.energy = energy[p];
    S.epre = epre ? epre[p] : 0.0;
    const double* f = fs + FS_NF * p;
    S.fsScale = f[FS_SCALE];
    S.fsOneMa = f[FS_ONEMA];
    S.fsMaxF = f[FS_MAXF];
    S.fsMaxFB = f[FS_MAXFB];
    return S;
}
__device__ __forceinline__ void fire_store(int p, const FireRegs& S, int* active, int* flags, double* dt,
                                           double* alpha, int* nPos, int* steps, double* energy, double* fs) {
    active[p] = S.act;
#pragma unroll
    for (int k = 0; k < FL_NF; ++k) flags[FL_NF * p + k] = S.fl[k];
    dt[p] = S.dt;
    alpha[p] = S.alpha;
    nPos[p] = S.nPos;
    steps[p] = S.steps;
    energy[p] = S.energy;
    double* f = fs + FS_NF * p;
    f[FS_SCALE] = S.fsScale;
    f[FS_ONEMA] = S.fsOneMa;
    f[FS_MAXF] = S.fsMaxF;
}

// energy test (model.cpp:444-470) on the packing's new energy S.energy
__device__ __forceinline__ void fire_control(FireRegs& S, double alphaStart, double fDec) {
    if (!S.act) return;
    double ePost = S.energy, ePre = S.epre;
    if (S.fl[FL_FAIL]) {
        S.dt = S.dt * fDec;
        S.alpha = alphaStart;
        S.nPos = 0;
        S.fl[FL_ACCEPT] = 0;
    } else if (ePost > ePre + fmax(fabs(ePre) * 1e-10, 1e-14)) {
        S.fl[FL_ROLLBACK] = 1;
        S.fl[FL_ACCEPT] = 0;
        S.dt = S.dt * fDec;
        S.alpha = alphaStart;
        S.nPos = 0;
    } else {
        S.fl[FL_ACCEPT] = 1;
    }
}

// block per packing: energy = fixed-order sum of the polygon energies, then the test
__global__ __launch_bounds__(REDUCE_BLK) void k_fire_energy(int N, const int* __restrict__ packPoly,
                                                    const double* __restrict__ epoly, double* __restrict__ energy,
                                                    int* __restrict__ active, const double* __restrict__ epre,
                                                    double alphaStart, double fDec, double* __restrict__ dt,
                                                    double* __restrict__ alpha, int* __restrict__ nPos,
                                                    int* __restrict__ flags, int* __restrict__ steps,
                                                    double* __restrict__ fs, const int* __restrict__ xpbdIt,
                                                    int keepDone) {
    __shared__ double part[REDUCE_BLK + 1];
    int p = blockIdx.x;
    // keepDone: the evaluation skipped the converged packings, whose energy stays as FIRE left it
    if (keepDone && !active[p]) return;   // the whole workgroup
    int a0 = packPoly[p], n = packPoly[p + 1] - a0;
    FireRegs S;
    int xit = 0;
    if (threadIdx.x == 0) {
        S = fire_load(p, active, flags, dt, alpha, nPos, steps, energy, epre, fs);
        if (xpbdIt) xit = xpbdIt[p];
    }
    double e = det_sum_block<REDUCE_BLK>(epoly + a0, n, part);
    if (threadIdx.x == 0) {
        S.energy = e;
        if (xpbdIt && S.act) S.fl[FL_LASTSHAKE] = xit;   // XPBD iterations
        fire_control(S, alphaStart, fDec);
        fire_store(p, S, active, flags, dt, alpha, nPos, steps, energy, fs);
    }
}

// second half-kick and per-polygon partial sums of F.v, v.v, F.F (vertex order); max |F|
// NOKICK (FIRE with XPBD): no second half-kick, the sums only
template <bool NOKICK>
__global__ void k_fire_kick(int N, const int* __restrict__ polyStart, const int* __restrict__ polyPack,
                            const int* __restrict__ active, const int* __restrict__ flags,
                            const double* __restrict__ dt, double2* __restrict__ v,
                            const double2* __restrict__ F, double* __restrict__ red) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N) return;
    int p = polyPack[s];
    double a = 0.0, b = 0.0, c = 0.0, mf = 0.0;
    for (int m = polyStart[s]; m < polyStart[s + 1]; ++m) {   // max |force component|
        double2 fi = F[m];
        mf = fmax(mf, fmax(fabs(fi.x), fabs(fi.y)));
    }
    if (active[p] && flags[FL_NF * p + FL_ACCEPT]) {
        double h = dt[p];
        for (int m = polyStart[s]; m < polyStart[s + 1]; ++m) {
            double2 vi = v[m], fi = F[m];
            double vx = vi.x, vy = vi.y;
            if (!NOKICK) {
                vx = vi.x + (0.5 * h) * fi.x;
                vy = vi.y + (0.5 * h) * fi.y;
                v[m] = make_double2(vx, vy);
            }
            a += fi.x * vx;
            a += fi.y * vy;
            b += vx * vx;
            b += vy * vy;
            c += fi.x * fi.x;
            c += fi.y * fi.y;
        }
    }
    red[s] = a;
    red[N + s] = b;
    red[2LL * N + s] = c;
    red[3LL * N + s] = mf;
}

// k_fire_kick with one lane segment per polygon (n <= SEG): the vertices are read and
// written coalesced; the per-polygon sums keep the sequential vertex order (x then y term
// of vertex 0, of vertex 1, ...), replayed by each lane of the segment from the products
// staged in LDS.
template <int SEG, bool NOKICK>
__global__ __launch_bounds__(BLK) void k_fire_kick_seg(int N, const int* __restrict__ polyStart,
                                                      const int* __restrict__ polyPack,
                                                      const int* __restrict__ active, const int* __restrict__ flags,
                                                      const double* __restrict__ dt, double2* __restrict__ v,
                                                      const double2* __restrict__ F, double* __restrict__ red) {
    __shared__ double2 prod[3][BLK];   // (F.x v.x, F.y v.y), (v.x v.x, v.y v.y), (F.x F.x, F.y F.y)
    int e = threadIdx.x & (SEG - 1);
    int segBase = threadIdx.x - e;
    int s = (blockIdx.x * BLK + threadIdx.x) / SEG;
    int st = 0, n = 0, p = 0;
    if (s < N) {
        st = polyStart[s];
        n = polyStart[s + 1] - st;
        p = polyPack[s];
    }
    bool live = e < n;
    double2 fi = live ? F[st + e] : make_double2(0.0, 0.0);
    double mf = fmax(fabs(fi.x), fabs(fi.y));
#pragma unroll
    for (int o = SEG / 2; o >= 1; o >>= 1) mf = fmax(mf, __shfl_xor(mf, o, SEG));
    bool kick = (s < N) && active[p] && flags[FL_NF * p + FL_ACCEPT];
    double2 pa = make_double2(0.0, 0.0), pb = pa, pc = pa;
    if (kick && live) {
        double h = dt[p];
        double2 vi = v[st + e];
        double vx = vi.x, vy = vi.y;
        if (!NOKICK) {
            vx = vi.x + (0.5 * h) * fi.x;
            vy = vi.y + (0.5 * h) * fi.y;
            v[st + e] = make_double2(vx, vy);
        }
        pa = make_double2(fi.x * vx, fi.y * vy);
        pb = make_double2(vx * vx, vy * vy);
        pc = make_double2(fi.x * fi.x, fi.y * fi.y);
    }
    prod[0][threadIdx.x] = pa;
    prod[1][threadIdx.x] = pb;
    prod[2][threadIdx.x] = pc;
    __syncthreads();
    double a = 0.0, b = 0.0, c = 0.0;
    if (kick)
        for (int m = 0; m < n; ++m) {
            double2 qa = prod[0][segBase + m], qb = prod[1][segBase + m], qc = prod[2][segBase + m];
            a += qa.x;
            a += qa.y;
            b += qb.x;
            b += qb.y;
            c += qc.x;
            c += qc.y;
        }
    if (s < N && e == 0) {
        red[s] = a;
        red[N + s] = b;
        red[2LL * N + s] = c;
        red[3LL * N + s] = mf;
    }
}

// bend + P>0 logic (model.cpp:471-485)
__device__ __forceinline__ void fire_control2(FireRegs& S, double Pw, double vv, double ff, int nMin, double dtMax,
                                              double alphaStart, double fAlpha, double fInc, double fDec) {
    if (!S.act) return;
    if (!S.fl[FL_ACCEPT]) return;
    double vnorm = sqrt(vv), fnorm = sqrt(ff);
    double a = S.alpha;
    S.fsScale = (fnorm > 0.0) ? (a * vnorm) / fnorm : 0.0;
    S.fsOneMa = 1.0 - a;
    double h = S.dt;
    if (Pw > 0.0) {
        S.fl[FL_ZEROV] = 0;
        int np = S.nPos + 1;
        S.nPos = np;
        if (np >= nMin) {
            double newDt = fmin(h * fInc, dtMax);
            if (newDt > h) S.alpha = S.alpha * fAlpha;
            S.dt = newDt;
        }
    } else {
        S.fl[FL_ZEROV] = 1;
        S.dt = h * fDec;
        S.alpha = alphaStart;
        S.nPos = 0;
    }
}

__global__ void k_fire_bend(int V, const int* __restrict__ shapeId, const int* __restrict__ polyPack,
                            const int* __restrict__ active, const int* __restrict__ flags,
                            const double* __restrict__ fs, double2* __restrict__ x, double2* __restrict__ v,
                            double2* __restrict__ F, const double2* __restrict__ xbak,
                            const double2* __restrict__ fbak, double2* __restrict__ Fe,
                            const double2* __restrict__ febak) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    int p = polyPack[shapeId[i]];
    const int* fl = flags + FL_NF * p;
    if (!fl[FL_ACTIVEBAK]) return;   // active when the step began
    if (fl[FL_ACCEPT]) {
        if (fl[FL_ZEROV]) {
            v[i] = make_double2(0.0, 0.0);
        } else {
            double sc = fs[FS_NF * p + FS_SCALE], om = fs[FS_NF * p + FS_ONEMA];
            double2 vi = v[i], fi = F[i];
            v[i] = make_double2(om * vi.x + sc * fi.x, om * vi.y + sc * fi.y);
        }
    } else if (fl[FL_ROLLBACK]) {
        x[i] = xbak[i];
        v[i] = make_double2(0.0, 0.0);
        if (fbak) F[i] = fbak[i];
        if (febak) Fe[i] = febak[i];
    }
}

// block-start snapshot of the FIRE state (save = 1) or its restoration (save = 0)
__global__ void k_fire_snapshot(int V, int P, int save, double2* __restrict__ x, double2* __restrict__ v,
                                double2* __restrict__ F, double* __restrict__ dt, double* __restrict__ alpha,
                                double* __restrict__ energy, double* __restrict__ fs, int* __restrict__ nPos,
                                int* __restrict__ steps, int* __restrict__ active, double2* __restrict__ sx,
                                double2* __restrict__ sv, double2* __restrict__ sf, double* __restrict__ sd,
                                int* __restrict__ si) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < V) {
        if (save) {
            sx[i] = x[i]; sv[i] = v[i]; sf[i] = F[i];
        } else {
            x[i] = sx[i]; v[i] = sv[i]; F[i] = sf[i];
        }
    }
    if (i < P) {
        double* d = sd + 4LL * i;
        int* k = si + 3LL * i;
        if (save) {
            d[0] = dt[i]; d[1] = alpha[i]; d[2] = energy[i]; d[3] = fs[FS_NF * i + FS_MAXF];
            k[0] = nPos[i]; k[1] = steps[i]; k[2] = active[i];
        } else {
            dt[i] = d[0]; alpha[i] = d[1]; energy[i] = d[2]; fs[FS_NF * i + FS_MAXF] = d[3];
            nPos[i] = k[0]; steps[i] = k[1]; active[i] = k[2];
        }
    }
}

// convergence on max |force component| (model.cpp:500), step count, status block.
// staleRollback: a rolled-back packing's forces are not yet recomputed, so it cannot
// converge in this pass.
__device__ __forceinline__ void fire_finish(int p, FireRegs& S, double mf, double fTol,
                                            pcp_fire_status* __restrict__ status, int* __restrict__ summary,
                                            int countStep, int staleRollback) {
    if (S.act) {
        S.fsMaxF = mf;
        if (countStep) S.steps = S.steps + 1;
        bool conv = (mf <= fTol) && !(staleRollback && S.fl[FL_ROLLBACK]);
        if (conv) S.act = 0;
        if (S.fl[FL_ROLLBACK]) atomicOr(&summary[0], 1);
        if (!conv) atomicOr(&summary[1], 1);
    }
    pcp_fire_status st;
    st.dt = S.dt;
    st.alpha = S.alpha;
    st.energy = S.energy;
    st.maxForce = S.fsMaxF;
    st.nPos = S.nPos;
    st.steps = S.steps;
    st.converged = S.act ? 0 : 1;
    st.lastShakeIters = S.fl[FL_LASTSHAKE];
    st.shakeFailed = S.fl[FL_FAIL];
    st.rolledBack = S.fl[FL_ROLLBACK];
    status[p] = st;
}

// Block per packing, after the kick: fixed-order sums of F.v, v.v, F.F and the max
// |force component|, then (full) the dt / alpha update (model.cpp:471-485) and the
// convergence test, or (full = 0, after a rollback's fresh forces) the test only.
// Block 0 also reports the largest record bucket count of the step in summary[2].
__global__ __launch_bounds__(REDUCE_BLK) void k_fire_stats(int N, const int* __restrict__ packPoly,
                                                   const double* __restrict__ red, int full, int nMin,
                                                   double dtMax, double alphaStart, double fAlpha, double fInc,
                                                   double fDec, double fTol, int* __restrict__ active,
                                                   int* __restrict__ flags, double* __restrict__ dt,
                                                   double* __restrict__ alpha, int* __restrict__ nPos,
                                                   int* __restrict__ steps, double* __restrict__ energy,
                                                   double* __restrict__ fs, pcp_fire_status* __restrict__ status,
                                                   int* __restrict__ summary, const int* __restrict__ scal,
                                                   const double* __restrict__ epre, int* __restrict__ blockOver) {
    __shared__ double L[REDUCE_BLK / 64];
    __shared__ double part[3 * (REDUCE_BLK + 1)];
    int p = blockIdx.x;
    int a0 = packPoly[p], n = packPoly[p + 1] - a0;
    FireRegs S;
    if (threadIdx.x == 0) S = fire_load(p, active, flags, dt, alpha, nPos, steps, energy, epre, fs);
    double Pw = 0.0, vv = 0.0, ff = 0.0, mf;
    if (full) {   // F.v, v.v, F.F and max |F| in one pass
        double r[4];
        det_reduce_block<3, true, REDUCE_BLK>(red + a0, N, n, part, L, r);
        Pw = r[0]; vv = r[1]; ff = r[2]; mf = r[3];
    } else {
        double r[1];
        det_reduce_block<0, true, REDUCE_BLK>(red + 3LL * N + a0, N, n, part, L, r);
        mf = r[0];
    }
    if (threadIdx.x == 0) {
        if (full) fire_control2(S, Pw, vv, ff, nMin, dtMax, alphaStart, fAlpha, fInc, fDec);
        // device-side rollback (epre given): the step's start state comes back from its
        // backups -- energy here, x, v and F in k_fire_bend -- as re-evaluating it would give
        if (full && epre && S.act && S.fl[FL_ROLLBACK]) {
            S.energy = S.epre;
            mf = S.fsMaxFB;
        }
        fire_finish(p, S, mf, fTol, status, summary, full, full);
        fire_store(p, S, active, flags, dt, alpha, nPos, steps, energy, fs);
        if (p == 0) {
            summary[2] = scal[SC_BUCKET];
            if (blockOver && scal[SC_BUCKET] > blockOver[0]) blockOver[0] = scal[SC_BUCKET];
        }
    }
}

// One step's FIRE bookkeeping for a context of a few small packings, one workgroup per packing:
// the energy test (k_fire_energy), the second half-kick with the per-polygon sums
// (a thread per polygon), the packing's sums, the dt / alpha update and
// the convergence test (k_fire_stats, full) and the bend or rollback of every vertex
// (k_fire_bend), with workgroup barriers where the four launches had kernel boundaries.  The same
// functions in the same order, so the results are those of the four launches bit for bit; a
// small context's step is bound by launch and latency, not by the work of one packing.
constexpr int FIRE_SMALL_B = 8;
template <bool NOKICK>
__global__ __launch_bounds__(REDUCE_BLK) void k_fire_small(
    int N, const int* __restrict__ packPoly, const int* __restrict__ packVert, const int* __restrict__ polyStart,
    const double* __restrict__ epoly, double* __restrict__ energy, int* __restrict__ active,
    const double* __restrict__ epre, const int* __restrict__ xpbdIt, int nMin, double dtMax, double alphaStart,
    double fAlpha, double fInc, double fDec, double fTol, double* __restrict__ dt, double* __restrict__ alpha,
    int* __restrict__ nPos, int* __restrict__ flags, int* __restrict__ steps, double* __restrict__ fs,
    pcp_fire_status* __restrict__ status,
    int* __restrict__ summary, const int* __restrict__ scal, int rollbackOnDevice, int* __restrict__ blockOver,
    double* __restrict__ red, double2* __restrict__ x, double2* __restrict__ v, double2* __restrict__ F,
    const double2* __restrict__ xbak, const double2* __restrict__ fbak, double2* __restrict__ Fe,
    const double2* __restrict__ febak, int keepDone) {
    __shared__ double part[3 * (REDUCE_BLK + 1)];
    __shared__ double L[REDUCE_BLK / 64];
    int p = blockIdx.x;
    int a0 = packPoly[p], n = packPoly[p + 1] - a0;
    // thread 0 holds the packing's FIRE scalars from here to the stats (written back before each
    // barrier after which other threads read them)
    FireRegs S;
    int xit = 0;
    if (threadIdx.x == 0) {
        S = fire_load(p, active, flags, dt, alpha, nPos, steps, energy, epre, fs);
        if (xpbdIt) xit = xpbdIt[p];
    }
    // energy = fixed-order sum of the polygon energies, then its test (keepDone: as k_fire_energy)
    double en = det_sum_block<REDUCE_BLK>(epoly + a0, n, part);
    if (threadIdx.x == 0) {
        if (!keepDone || S.act) S.energy = en;
        if (xpbdIt && S.act) S.fl[FL_LASTSHAKE] = xit;   // XPBD iterations
        fire_control(S, alphaStart, fDec);
        fire_store(p, S, active, flags, dt, alpha, nPos, steps, energy, fs);
    }
    __syncthreads();
    // second half-kick and per-polygon sums in vertex order (as k_fire_kick / k_fire_kick_seg), a
    // thread per polygon, its vertices loaded FIRE_SMALL_B at a time with the loads in flight together
    bool kickP = active[p] && flags[FL_NF * p + FL_ACCEPT];
    double h = dt[p];
    for (int s = a0 + threadIdx.x; s < a0 + n; s += REDUCE_BLK) {
        int st = polyStart[s], nv = polyStart[s + 1] - st;
        double a = 0.0, b = 0.0, c = 0.0, mf = 0.0;
        for (int m0 = 0; m0 < nv; m0 += FIRE_SMALL_B) {
            double2 fq[FIRE_SMALL_B], vq[FIRE_SMALL_B];
#pragma unroll
            for (int k = 0; k < FIRE_SMALL_B; ++k) {
                bool in = m0 + k < nv;
                fq[k] = in ? F[st + m0 + k] : make_double2(0.0, 0.0);
                vq[k] = (in && kickP) ? v[st + m0 + k] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int k = 0; k < FIRE_SMALL_B; ++k) {
                if (m0 + k >= nv) break;
                double2 fi = fq[k];
                mf = fmax(mf, fmax(fabs(fi.x), fabs(fi.y)));
                if (kickP) {
                    double vx = vq[k].x, vy = vq[k].y;
                    if (!NOKICK) {
                        vx = vq[k].x + (0.5 * h) * fi.x;
                        vy = vq[k].y + (0.5 * h) * fi.y;
                        v[st + m0 + k] = make_double2(vx, vy);
                    }
                    a += fi.x * vx;
                    a += fi.y * vy;
                    b += vx * vx;
                    b += vy * vy;
                    c += fi.x * fi.x;
                    c += fi.y * fi.y;
                }
            }
        }
        red[s] = a;
        red[N + s] = b;
        red[2LL * N + s] = c;
        red[3LL * N + s] = mf;
    }
    __syncthreads();
    // the packing's sums of F.v, v.v, F.F and max |F|, dt / alpha, convergence, as k_fire_stats
    double r[4];
    det_reduce_block<3, true, REDUCE_BLK>(red + a0, N, n, part, L, r);
    if (threadIdx.x == 0) {
        fire_control2(S, r[0], r[1], r[2], nMin, dtMax, alphaStart, fAlpha, fInc, fDec);
        double mfp = r[3];
        if (rollbackOnDevice && S.act && S.fl[FL_ROLLBACK]) {
            S.energy = S.epre;
            mfp = S.fsMaxFB;
        }
        fire_finish(p, S, mfp, fTol, status, summary, 1, 1);
        fire_store(p, S, active, flags, dt, alpha, nPos, steps, energy, fs);
        if (p == 0) {
            summary[2] = scal[SC_BUCKET];
            if (blockOver && scal[SC_BUCKET] > blockOver[0]) blockOver[0] = scal[SC_BUCKET];
        }
    }
    __syncthreads();
    // bend or rollback of every vertex of the packing, as k_fire_bend (FIRE_SMALL_B vertices per
    // thread at a time, their loads in flight together)
    const int* fl = flags + FL_NF * p;
    if (fl[FL_ACTIVEBAK]) {
        bool accept = fl[FL_ACCEPT], zero = fl[FL_ZEROV], roll = fl[FL_ROLLBACK];
        double sc = fs[FS_NF * p + FS_SCALE], om = fs[FS_NF * p + FS_ONEMA];
        int v0 = packVert[p], v1 = packVert[p + 1];
        for (int i0 = v0 + threadIdx.x; i0 < v1; i0 += REDUCE_BLK * FIRE_SMALL_B) {
            if (accept && !zero) {
                double2 vq[FIRE_SMALL_B], fq[FIRE_SMALL_B];
#pragma unroll
                for (int k = 0; k < FIRE_SMALL_B; ++k) {
                    int i = i0 + k * REDUCE_BLK;
                    vq[k] = (i < v1) ? v[i] : make_double2(0.0, 0.0);
                    fq[k] = (i < v1) ? F[i] : make_double2(0.0, 0.0);
                }
#pragma unroll
                for (int k = 0; k < FIRE_SMALL_B; ++k) {
                    int i = i0 + k * REDUCE_BLK;
                    if (i < v1)
                        v[i] = make_double2(om * vq[k].x + sc * fq[k].x, om * vq[k].y + sc * fq[k].y);
                }
            } else if (accept) {
#pragma unroll
                for (int k = 0; k < FIRE_SMALL_B; ++k) {
                    int i = i0 + k * REDUCE_BLK;
                    if (i < v1) v[i] = make_double2(0.0, 0.0);
                }
            } else if (roll) {
                double2 xq[FIRE_SMALL_B], fq[FIRE_SMALL_B], eq[FIRE_SMALL_B];
#pragma unroll
                for (int k = 0; k < FIRE_SMALL_B; ++k) {
                    int i = i0 + k * REDUCE_BLK;
                    bool in = i < v1;
                    xq[k] = in ? xbak[i] : make_double2(0.0, 0.0);
                    fq[k] = (in && fbak) ? fbak[i] : make_double2(0.0, 0.0);
                    eq[k] = (in && febak) ? febak[i] : make_double2(0.0, 0.0);
                }
#pragma unroll
                for (int k = 0; k < FIRE_SMALL_B; ++k) {
                    int i = i0 + k * REDUCE_BLK;
                    if (i < v1) {
                        x[i] = xq[k];
                        v[i] = make_double2(0.0, 0.0);
                        if (fbak) F[i] = fq[k];
                        if (febak) Fe[i] = eq[k];
                    }
                }
            }
        }
    }
}

// undo an attempt whose buffers overflowed: x, v, FIRE scalars, active flags, steps
__global__ void k_fire_undo(int V, int P, const int* __restrict__ shapeId, const int* __restrict__ polyPack,
                            int* __restrict__ active, const int* __restrict__ flags, double2* __restrict__ x,
                            double2* __restrict__ v, const double2* __restrict__ xbak,
                            const double2* __restrict__ vbak, double* dt, double* alpha, int* nPos,
                            int* steps, const double* __restrict__ fs) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < V) {
        int p = polyPack[shapeId[i]];
        if (flags[FL_NF * p + FL_ACTIVEBAK]) { x[i] = xbak[i]; v[i] = vbak[i]; }
    }
    if (i < P && flags[FL_NF * i + FL_ACTIVEBAK]) {
        active[i] = 1;
        dt[i] = fs[FS_NF * i + FS_DTB];
        alpha[i] = fs[FS_NF * i + FS_ALB];
        nPos[i] = (int)fs[FS_NF * i + FS_NPB];
        steps[i] = (int)fs[FS_NF * i + FS_STEPB];
    }
}

__global__ void k_zero2(long long n, double2* v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = make_double2(0.0, 0.0);
}

// the evaluation part of one step (positions already moved by k_fire_begin)
// the kick: lane segments when every polygon fits one (n <= 64), else thread per polygon
template <bool NOKICK>
static int launch_fire_kick(pcp_ctx* c, double2* vel, const double2* force) {
    int seg = c->seg;
    if (c->maxN > 64) {
        PCP_LAUNCH(c, k_fire_kick<NOKICK>, dim3(nblk(c->N)), dim3(BLK), c->N, c->dPolyStart, c->dPolyPack,
                   c->dActive, c->dFlags, c->dDt, vel, force, c->dRed);
        return PCP_OK;
    }
    int grid = (c->N + BLK / seg - 1) / (BLK / seg);
#define PCP_KICK_SEG(W)                                                                                     \
    PCP_LAUNCH(c, (k_fire_kick_seg<W, NOKICK>), dim3(grid), dim3(BLK), c->N, c->dPolyStart, c->dPolyPack, \
               c->dActive, c->dFlags, c->dDt, vel, force, c->dRed)
    switch (seg) {
    case 4: PCP_KICK_SEG(4); break;
    case 8: PCP_KICK_SEG(8); break;
    case 16: PCP_KICK_SEG(16); break;
    case 32: PCP_KICK_SEG(32); break;
    default: PCP_KICK_SEG(64); break;
    }
#undef PCP_KICK_SEG
    return PCP_OK;
}

// A context of at least SKIP_DONE_MIN_V vertices in several packings is throughput-bound: its FIRE
// evaluations skip the packings that have converged (their polygons are left out of the
// detection and record kernels' work list).  Below that, a step is bound by its kernels'
// dependent-load chains, which the extra load of the list's length would lengthen.
constexpr int SKIP_DONE_MIN_V = 1 << 17;
static bool skip_done(const pcp_ctx* c) { return c->P > 1 && c->V >= SKIP_DONE_MIN_V; }
struct SkipDoneScope {   // c->skipDone for the launches of one evaluation
    pcp_ctx* c;
    SkipDoneScope(pcp_ctx* ctx, bool on) : c(ctx) { c->skipDone = on; }
    ~SkipDoneScope() { c->skipDone = false; }
};

// a context of at most FIRE_SMALL_P packings of at most FIRE_SMALL_V vertices each, polygons of
// at most 64 vertices: its FIRE bookkeeping runs as one launch (k_fire_small)
constexpr int FIRE_SMALL_P = 16;
constexpr int FIRE_SMALL_V = 32768;
static bool fire_small(const pcp_ctx* c) {
    if (c->P > FIRE_SMALL_P || c->maxN > 64) return false;
    for (int p = 0; p < c->P; ++p)
        if (c->hPackVert[p + 1] - c->hPackVert[p] > FIRE_SMALL_V) return false;
    return true;
}
template <bool NOKICK>
static int launch_fire_small(pcp_ctx* c, const pcp_state* st, const pcp_fire_params* prm, int block,
                             const double2* fbak, const double2* febak) {
PCP_LAUNCH(c, (k_fire_small<NOKICK>), dim3(c->P), dim3(REDUCE_BLK), c->N, c->dPackPoly, c->dPackVert, c->dPolyStart,
               c->dEpoly, st->energy, c->dActive, c->dEpre, NOKICK ? (const int*)c->dShakeIt : (const int*)nullptr,
               prm->nMin, prm->dtMax, prm->alphaStart, prm->fAlpha, prm->fInc, prm->fDec, prm->fTol, c->dDt,
               c->dAlpha, c->dNPos, c->dFlags, c->dSteps, c->dFireScal, c->dStatus, c->dSummary, c->dScal, block,
               block ? c->dBlockOver : nullptr, c->dRed, (double2*)st->pos, (double2*)st->vel,
               (double2*)st->force, (const double2*)c->dXbak, fbak, (double2*)c->dFext, febak, skip_done(c) ? 1 : 0);
    return PCP_OK;
}

// energyTest: end with the energy test (k_fire_energy); a small context does it in k_fire_small
static int fire_eval(pcp_ctx* c, const pcp_state* st, const pcp_fire_params* prm, bool energyTest = true) {
    double2* x = (double2*)st->pos;
    double2* v = (double2*)st->vel;
    int mt = st->model & 0xFF;
    bool area = (st->model & PCP_AREA_MODES) != 0;
    bool overlap = (mt == PCP_NORMAL || mt == PCP_HYBRID);
    // the detection's cell sort lists the running packings' polygons (models with overlap)
    SkipDoneScope skipScope(c, overlap && skip_done(c));
    bool rigid = (mt == PCP_NORMAL);
    bool xpbd = prm->xpbdIter > 0;
    bool cons = (st->model & PCP_CONSISTENT) != 0;
    if (xpbd) {
        // the springs of the model by XPBD at the packing's dt, then velocities from the
        // displacement; detection and records at the final positions
        bool edgeOn = (mt == PCP_SOFT_BODY || mt == PCP_EDGE_ONLY || mt == PCP_HYBRID);
        bool areaOn = (mt == PCP_SOFT_BODY || mt == PCP_AREA_ONLY || mt == PCP_HYBRID);
        bool perimOn = (mt == PCP_SOFT_BODY || mt == PCP_HYBRID);
        PCP_TRY(run_xpbd_seg(c, x, st->targetEdgeV, nullptr, st->targetAreas, edgeOn ? st->stiffness : 0.0,
                             areaOn ? st->compressibility : 0.0, perimOn ? st->perimStiffness : 0.0, 0.0, c->dDt,
                             prm->xpbdIter, prm->xpbdTol, c->dActive, v, c->dShakeIt));
        if (overlap) {
            PCP_TRY(run_detect(c, x, 0, nullptr, cons));
            PCP_TRY(run_records(c, area ? x : nullptr, true));
        }
    } else if (overlap) {
        if (!area)   // reference order (model.cpp:416-440): detect, SHAKE, records
            PCP_TRY(run_detect(c, x, 0, nullptr, cons));
        bool shake = rigid && prm->shakeIter > 0;
        if (shake)
            PCP_TRY(run_shake(c, x, st->targetEdgeV, st->targetAreas, prm->shakeIter, 1e-15,
                              10 * prm->shakeIter, c->dActive, c->dShakeIt));
        if (shake && area) {
            // area energies: crossings of the final positions; the restore-or-rederive step
            // after SHAKE is done by the detection's first kernel (k_bounds<POST>)
            PostShake ps{prm->shakeIter, FL_NF, FL_FAIL, FL_LASTSHAKE, c->dPolyPack, c->dPackVert, c->dActive,
                         c->dDt, c->dShakeIt, x, v, (const double2*)c->dXbak, (const double2*)c->dXpre, c->dFlags};
            PCP_TRY(run_detect(c, x, 0, &ps, cons));
        } else {
            if (shake)
                PCP_LAUNCH(c, k_fire_post_shake, dim3(nblk(c->V)), dim3(BLK), c->V, c->P,
                           prm->shakeIter, c->dShapeId, c->dPolyPack, c->dPackVert, c->dActive, c->dDt,
                           c->dShakeIt, x, v, (const double2*)c->dXbak, (const double2*)c->dXpre,
                           c->dFlags);
            if (area)   // area energy: crossings of the final positions
                PCP_TRY(run_detect(c, x, 0, nullptr, cons));
        }
        PCP_TRY(run_records(c, area ? x : nullptr, true));
    }
    PCP_TRY(run_forces(c, st->model, x, st->edgeLengths, st->areas, st->targetEdgeV, st->targetAreas,
                       st->stiffness, st->compressibility, (double2*)st->force, nullptr, 1, rigid, false,
                       st->perimStiffness));
    if (!energyTest) return PCP_OK;
    PCP_LAUNCH(c, k_fire_energy, dim3(c->P), dim3(REDUCE_BLK), c->N, c->dPackPoly, c->dEpoly, st->energy, c->dActive,
               c->dEpre, prm->alphaStart, prm->fDec, c->dDt, c->dAlpha, c->dNPos, c->dFlags, c->dSteps,
               c->dFireScal, xpbd ? (const int*)c->dShakeIt : (const int*)nullptr, skip_done(c) ? 1 : 0);
    return PCP_OK;
}

// read the status block (summary[4] then the per-packing status) back to the host
static int fire_read_status(pcp_ctx* c) {
    size_t bytes = 16 + sizeof(pcp_fire_status) * (size_t)c->P;
    c->hStatusBlock.resize(bytes);
    PCP_HIP(hipMemcpyAsync(c->hStatusBlock.data(), c->dStatusBlock, bytes, hipMemcpyDeviceToHost, c->stream));
    PCP_HIP(hipStreamSynchronize(c->stream));
    return PCP_OK;
}
static const int* host_summary(pcp_ctx* c) { return (const int*)c->hStatusBlock.data(); }

// Launches of one FIRE step (no host interaction).  block = 1: device-side rollback (the
// step's start state comes back from its backups) and the block's largest bucket count is
// kept in dBlockOver, so several steps can be queued without reading the status back.
static int fire_step_launch(pcp_ctx* c, const pcp_state* st, const pcp_fire_params* prm, int block) {
    double2* fbak = block ? (double2*)c->dFbak : nullptr;
    bool xpbd = prm->xpbdIter > 0;
    double2* febak = (block && xpbd) ? (double2*)c->dFextBak : nullptr;
    if (xpbd)
        PCP_LAUNCH(c, k_fire_begin<true>, dim3(nblk(std::max(std::max(c->V, c->P), 4))), dim3(BLK), c->V, c->P,
                   c->dShapeId, c->dPolyPack, c->dActive, c->dDt, c->dAlpha, c->dNPos, c->dSteps, st->energy,
                   c->dEpre, c->dFlags, c->dFireScal, c->dShakeIt, c->dSummary, (double2*)st->pos,
                   (double2*)st->vel, (const double2*)st->force, (double2*)c->dXbak, (double2*)c->dVbak,
                   (double2*)c->dXpre, fbak, (const double2*)c->dFext, febak);
    else
        PCP_LAUNCH(c, k_fire_begin<false>, dim3(nblk(std::max(std::max(c->V, c->P), 4))), dim3(BLK), c->V, c->P,
                   c->dShapeId, c->dPolyPack, c->dActive, c->dDt, c->dAlpha, c->dNPos, c->dSteps, st->energy,
                   c->dEpre, c->dFlags, c->dFireScal, c->dShakeIt, c->dSummary, (double2*)st->pos,
                   (double2*)st->vel, (const double2*)st->force, (double2*)c->dXbak, (double2*)c->dVbak,
                   (double2*)c->dXpre, fbak, (const double2*)nullptr, (double2*)nullptr);
    if (fire_small(c)) {   // energy test, kick, stats and bend in one launch
        PCP_TRY(fire_eval(c, st, prm, false));
        if (xpbd) return launch_fire_small<true>(c, st, prm, block, fbak, febak);
        return launch_fire_small<false>(c, st, prm, block, fbak, febak);
    }
    PCP_TRY(fire_eval(c, st, prm));
    if (xpbd) PCP_TRY(launch_fire_kick<true>(c, (double2*)st->vel, (const double2*)st->force));
    else PCP_TRY(launch_fire_kick<false>(c, (double2*)st->vel, (const double2*)st->force));
    PCP_LAUNCH(c, k_fire_stats, dim3(c->P), dim3(REDUCE_BLK), c->N, c->dPackPoly, c->dRed, 1, prm->nMin, prm->dtMax,
               prm->alphaStart, prm->fAlpha, prm->fInc, prm->fDec, prm->fTol, c->dActive, c->dFlags, c->dDt, c->dAlpha,
               c->dNPos, c->dSteps, st->energy, c->dFireScal, c->dStatus, c->dSummary, c->dScal,
               block ? (const double*)c->dEpre : nullptr, block ? c->dBlockOver : nullptr);
    PCP_LAUNCH(c, k_fire_bend, dim3(nblk(c->V)), dim3(BLK), c->V, c->dShapeId, c->dPolyPack, c->dActive, c->dFlags,
               c->dFireScal, (double2*)st->pos, (double2*)st->vel, (double2*)st->force, (const double2*)c->dXbak,
               (const double2*)fbak, (double2*)c->dFext, (const double2*)febak);
    return PCP_OK;
}

// one step; *anyActive receives whether some packing is still unconverged
static int fire_step_impl(pcp_ctx* c, const pcp_state* st, const pcp_fire_params* prm, int* anyActive) {
    int mt0 = st->model & 0xFF;
    bool overlap = (mt0 == PCP_NORMAL || mt0 == PCP_HYBRID);
    for (int attempt = 0; attempt < 6; ++attempt) {
        PCP_TRY(fire_step_launch(c, st, prm, 0));
        PCP_TRY(fire_read_status(c));
        const int* summary = host_summary(c);
        if (overlap && summary[2] > c->Kp) {
            // the step ran with too small record buckets: grow, undo, retry
            int need = summary[2];
            hipLaunchKernelGGL(k_fire_undo, dim3(nblk(std::max(c->V, c->P))), dim3(BLK), 0, c->stream, c->V,
                               c->P, c->dShapeId, c->dPolyPack, c->dActive, c->dFlags, (double2*)st->pos,
                               (double2*)st->vel, (const double2*)c->dXbak, (const double2*)c->dVbak, c->dDt,
                               c->dAlpha, c->dNPos, c->dSteps, c->dFireScal);
            PCP_LAUNCHED();
            PCP_TRY(grow_buckets(c, need));
            PCP_TRY(pcp_force_eval(c, st));
            PCP_HIP(hipMemsetAsync(c->dSticky, 0, sizeof(int), c->stream));   // handled here
            continue;
        }
        if (summary[0]) {
            // energy rose for some packing: x was restored there; recompute its forces
            // (model.cpp:454-464) -- deterministic, so unchanged packings reproduce theirs
            PCP_TRY(pcp_force_eval(c, st));
            PCP_LAUNCH(c, k_max_abs_poly, dim3(nblk(c->N)), dim3(BLK), c->N, c->dPolyStart,
                       (const double2*)st->force, c->dRed + 3LL * c->N);
            PCP_LAUNCH(c, k_fire_stats, dim3(c->P), dim3(REDUCE_BLK), c->N, c->dPackPoly, c->dRed, 0, prm->nMin,
                       prm->dtMax, prm->alphaStart, prm->fAlpha, prm->fInc, prm->fDec, prm->fTol, c->dActive,
                       c->dFlags, c->dDt, c->dAlpha, c->dNPos, c->dSteps, st->energy, c->dFireScal, c->dStatus,
                       c->dSummary, c->dScal, (const double*)nullptr, (int*)nullptr);
            PCP_TRY(fire_read_status(c));
            summary = host_summary(c);
        }
        if (anyActive) *anyActive = summary[1];
        return PCP_OK;
    }
    set_error("pcp_fire_step: record buckets kept overflowing");
    return PCP_ERR_CAPACITY;
}

static int fire_snapshot(pcp_ctx* c, const pcp_state* st, int save) {
    PCP_LAUNCH(c, k_fire_snapshot, dim3(nblk(std::max(c->V, c->P))), dim3(BLK), c->V, c->P, save, (double2*)st->pos,
               (double2*)st->vel, (double2*)st->force, c->dDt, c->dAlpha, st->energy, c->dFireScal, c->dNPos,
               c->dSteps, c->dActive, (double2*)c->dSnapX, (double2*)c->dSnapV, (double2*)c->dSnapF, c->dSnapD,
               c->dSnapI);
    return PCP_OK;
}

// K steps queued back to back: replayed from a captured graph of the K-step launch sequence
// (captured again when K, the bucket capacity, the parameters or the state buffers change)
// or launched directly
static int fire_block_launch(pcp_ctx* c, const pcp_state* st, const pcp_fire_params* prm, int K, int useGraph) {
    if (!useGraph || !c->profName.empty()) {
        for (int k = 0; k < K; ++k) PCP_TRY(fire_step_launch(c, st, prm, 1));
        return PCP_OK;
    }
    bool valid = c->fireGraph && c->graphK == K && c->graphKp == c->Kp &&
                 memcmp(&c->graphPrm, prm, sizeof(*prm)) == 0 && memcmp(&c->graphSt, st, sizeof(*st)) == 0;
    if (!valid) {
        if (c->fireGraph) PCP_HIP(hipGraphExecDestroy(c->fireGraph));
        c->fireGraph = nullptr;
        PCP_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = PCP_OK;
        for (int k = 0; k < K && rc == PCP_OK; ++k) rc = fire_step_launch(c, st, prm, 1);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc != PCP_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail_hip(e, "hipStreamEndCapture", __FILE__, __LINE__);
        e = hipGraphInstantiate(&c->fireGraph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            c->fireGraph = nullptr;
            return fail_hip(e, "hipGraphInstantiate", __FILE__, __LINE__);
        }
        c->graphK = K;
        c->graphKp = c->Kp;
        c->graphPrm = *prm;
        c->graphSt = *st;
    }
    PCP_HIP(hipGraphLaunch(c->fireGraph, c->stream));
    return PCP_OK;
}

// one block of K steps with one status read; a block in which a record bucket overflowed is
// undone from the block-start snapshot and run again with larger buckets
static int fire_block(pcp_ctx* c, const pcp_state* st, const pcp_fire_params* prm, int K, int useGraph,
                      int* anyActive) {
    int mt0 = st->model & 0xFF;
    bool overlap = (mt0 == PCP_NORMAL || mt0 == PCP_HYBRID);
    for (int attempt = 0; attempt < 6; ++attempt) {
        PCP_TRY(fire_snapshot(c, st, 1));
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, c->dBlockOver, 1LL, 0);
        PCP_LAUNCHED();
        PCP_TRY(fire_block_launch(c, st, prm, K, useGraph));
        int over = 0;
        PCP_HIP(hipMemcpyAsync(&over, c->dBlockOver, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        PCP_TRY(fire_read_status(c));   // synchronizes
        const int* summary = host_summary(c);
        if (overlap && over > c->Kp) {
            PCP_TRY(fire_snapshot(c, st, 0));
            PCP_TRY(grow_buckets(c, over));
            PCP_TRY(pcp_force_eval(c, st));
            PCP_HIP(hipMemsetAsync(c->dSticky, 0, sizeof(int), c->stream));
            PCP_HIP(hipStreamSynchronize(c->stream));
            continue;
        }
        if (anyActive) *anyActive = summary[1];
        return PCP_OK;
    }
    set_error("pcp_fire_run: record buckets kept overflowing");
    return PCP_ERR_CAPACITY;
}

}  // namespace pcp

extern "C" {

static int check_fire_params(pcp_ctx* c, const pcp_state* st, const pcp_fire_params* prm) {
    int mt = st->model & 0xFF;
    if (prm->xpbdIter > 0 && (mt == PCP_NORMAL || mt == PCP_ABNORMAL)) {
        set_error("FIRE with XPBD (xpbdIter > 0) integrates the springs of a model that has them "
                  "(edgeOnly, areaOnly, softBody, hybrid)");
        return PCP_ERR_ARG;
    }
    if (prm->xpbdIter > 0 && c->maxN > 64) {
        set_error("FIRE with XPBD: polygons of at most 64 vertices");
        return PCP_ERR_ARG;
    }
    return PCP_OK;
}

int pcp_fire_init(pcp_ctx* c, const pcp_state* st, const pcp_fire_params* prm, double dt0) {
    PCP_TRY(check_fire_params(c, st, prm));
    if (!c->haveBoxes && ((st->model & 0xFF) == PCP_NORMAL || (st->model & 0xFF) == PCP_HYBRID)) {
        set_error("pcp_fire_init: set the box sizes first");
        return PCP_ERR_ARG;
    }
    hipLaunchKernelGGL(k_zero2, dim3(nblk(c->V)), dim3(BLK), 0, c->stream, (long long)c->V, (double2*)st->vel);
    PCP_LAUNCHED();
    hipLaunchKernelGGL(k_fire_init, dim3(nblk(c->P)), dim3(BLK), 0, c->stream, c->P, dt0, prm->alphaStart,
                       c->dDt, c->dAlpha, c->dNPos, c->dSteps, c->dActive, c->dFlags, c->dFireScal);
    PCP_LAUNCHED();
    PCP_TRY(pcp_force_eval(c, st));
    PCP_HIP(hipStreamSynchronize(c->stream));
    return PCP_OK;
}

int pcp_fire_step(pcp_ctx* c, const pcp_state* st, const pcp_fire_params* prm, pcp_fire_status* status) {
    PCP_TRY(check_fire_params(c, st, prm));
    PCP_TRY(fire_step_impl(c, st, prm, nullptr));
    if (status) memcpy(status, c->hStatusBlock.data() + 16, sizeof(pcp_fire_status) * c->P);
    return PCP_OK;
}

int pcp_fire_minimize(pcp_ctx* c, const pcp_state* st, const pcp_fire_params* prm, int maxSteps,
                      pcp_fire_status* status) {
    PCP_TRY(check_fire_params(c, st, prm));
    for (int step = 0; step < maxSteps; ++step) {
        int any = 1;
        PCP_TRY(fire_step_impl(c, st, prm, &any));
        if (!any) break;
    }
    if (status) {
        PCP_TRY(fire_read_status(c));
        memcpy(status, c->hStatusBlock.data() + 16, sizeof(pcp_fire_status) * c->P);
    }
    return PCP_OK;
}

int pcp_fire_run(pcp_ctx* c, const pcp_state* st, const pcp_fire_params* prm, int maxSteps, int stepsPerBlock,
                 int useGraph, pcp_fire_status* status) {
    PCP_TRY(check_fire_params(c, st, prm));
    if (stepsPerBlock < 1 || stepsPerBlock > 256) {
        set_error("pcp_fire_run: stepsPerBlock must be in 1..256");
        return PCP_ERR_ARG;
    }
    int done = 0;
    while (done < maxSteps) {
        int K = std::min(stepsPerBlock, maxSteps - done);
        int any = 1;
        PCP_TRY(fire_block(c, st, prm, K, useGraph && K == stepsPerBlock, &any));
        done += K;
        if (!any) break;
    }
    if (status) {
        PCP_TRY(fire_read_status(c));
        memcpy(status, c->hStatusBlock.data() + 16, sizeof(pcp_fire_status) * c->P);
    }
    return PCP_OK;
}

int pcp_fire_set(pcp_ctx* c, const double* dt, const double* alpha, const int* nPos) {
    PCP_HIP(hipStreamSynchronize(c->stream));
    if (dt) PCP_HIP(hipMemcpy(c->dDt, dt, sizeof(double) * c->P, hipMemcpyHostToDevice));
    if (alpha) PCP_HIP(hipMemcpy(c->dAlpha, alpha, sizeof(double) * c->P, hipMemcpyHostToDevice));
    if (nPos) PCP_HIP(hipMemcpy(c->dNPos, nPos, sizeof(int) * c->P, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(c->P)), dim3(BLK), 0, c->stream, c->dActive, (long long)c->P, 1);
    PCP_LAUNCHED();
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(c->P)), dim3(BLK), 0, c->stream, c->dSteps, (long long)c->P, 0);
    PCP_LAUNCHED();
    PCP_HIP(hipStreamSynchronize(c->stream));
    return PCP_OK;
}

int pcp_fire_get(pcp_ctx* c, pcp_fire_status* status) {
    PCP_HIP(hipStreamSynchronize(c->stream));
    PCP_HIP(hipMemcpy(status, c->dStatus, sizeof(pcp_fire_status) * c->P, hipMemcpyDeviceToHost));
    return PCP_OK;
}

}  // extern "C"
