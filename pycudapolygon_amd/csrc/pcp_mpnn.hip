// Learned warm start: a message-passing network over the contact graph of a packing.
// No counterpart in the reference (camdennis/pyCudaPolygon has no learned component);
// it sits in front of the plain method: its displacement is applied once, the polygons
// are projected back onto their shapes (SHAKE), and FP64 FIRE finishes the minimization
// and judges the residual force.
//
// Graph.  Nodes are vertices.  Vertex m (start of edge m) is linked to its polygon
// neighbours prev(m) and next(m), and to the start vertex of every edge that crosses
// edge m: each crossing record (si, sj, i, j) links vertex i of si with vertex j of sj,
// found from the si bucket of si (key_i == m) and the sj bucket of sj (key_j == m).
// Every crossing is one record, so the adjacency is symmetric.  It is stored with a
// fixed width of PCP_MPNN_DEG slots: prev, next, then the crossing partners in
// ascending global vertex id, -1 for empty slots (partners beyond the width are
// dropped and counted).
//
// Network (H = 64, parameters float32, layout PCP_MPNN_OFF_* in pcp.h):
//   h0  = relu(x W0 + b0)                                       x: 4 features
//   h_l = h_{l-1} + relu([h_{l-1}, sum_{u in N(m)} h_{l-1}(u)] W_l + b_l),  l = 1, 2, 3
//   y   = h3 Wo + bo                                            2 outputs per vertex
// Every product takes bf16 operands (weights and activations rounded to nearest even
// from float32) and accumulates in float32.  The 128x64 layers -- forward, input
// gradient and weight gradient -- run on the matrix cores (v_mfma_f32_16x16x32_bf16,
// one wavefront per 16 vertices); the 4x64 and 64x2 layers run on VALU in a fixed
// order.  Gradients treat the bf16 rounding as the identity (straight-through).
// Reductions over vertices are partial sums per chunk of vertices (MPNN_CH for the
// matrix-core weight gradients, MPNN_CHV for the VALU ones), then sums of groups of
// MPNN_GRP consecutive chunks, then the sum of the groups, each in order, so training is
// bit-reproducible run to run.
//
// This file is compiled as part of pcp_kernels.hip (one translation unit).

namespace pcp {

typedef short mpnn_s4 __attribute__((ext_vector_type(4)));
typedef short mpnn_s8 __attribute__((ext_vector_type(8)));
typedef __bf16 mpnn_b8 __attribute__((ext_vector_type(8)));
typedef float mpnn_f4 __attribute__((ext_vector_type(4)));

// One v_mfma_f32_16x16x32_bf16 step: lane (r, g = lane / 16) gives 8 bf16 of row r of A and 8
// of column r of B, element e of both at the same position k of the 32-long inner dimension
// (the hardware pairs element e of lane group g of A with element e of lane group g of B).
// Here the 8 elements are two groups of 4 (lo, hi) in the layout of the 16-wide steps: lane g
// holds k = 4g..4g+3 of two 16-wide halves, so one 32-wide step covers the same products as
// two 16-wide ones, in another addition order.
__device__ __forceinline__ mpnn_s8 cat_s4(mpnn_s4 lo, mpnn_s4 hi) {
    return (mpnn_s8){lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
}
__device__ __forceinline__ mpnn_f4 mfma32(mpnn_s8 a, mpnn_s8 b, mpnn_f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(mpnn_b8, a), __builtin_bit_cast(mpnn_b8, b), c,
                                                   0, 0, 0);
}

constexpr int MH = PCP_MPNN_H;           // hidden width
constexpr int MD = PCP_MPNN_DEG;         // adjacency slots per vertex
constexpr int MPNN_CH = 2048;            // vertices per partial-sum chunk (MFMA weight gradients)
constexpr int MPNN_CHV = 256;            // vertices per partial-sum chunk (VALU weight gradients)
constexpr int MPNN_GRP = 64;             // chunks per group in the sum of the chunk partials
constexpr int WPAD = 4;                  // LDS row padding (bf16 elements) against bank conflicts

__device__ __forceinline__ unsigned short bf16_rne(float f) {
    unsigned int u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_f(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }
__device__ __forceinline__ float bfr(float f) { return bf16_f(bf16_rne(f)); }

__device__ __forceinline__ mpnn_s4 pack_bf16(float a, float b, float c, float d) {
    mpnn_s4 r;
    r.x = (short)bf16_rne(a);
    r.y = (short)bf16_rne(b);
    r.z = (short)bf16_rne(c);
    r.w = (short)bf16_rne(d);
    return r;
}

// ---- contact graph -------------------------------------------------------------------
__global__ void k_mpnn_graph(int V, int Kp, const int* __restrict__ polyStart, const int* __restrict__ shapeId,
                             const int* __restrict__ polyPack, const int* __restrict__ packPoly,
                             const int* __restrict__ bcnt, const uint64_t* __restrict__ bkey,
                             const int* __restrict__ jcnt, const uint64_t* __restrict__ jkey,
                             int* __restrict__ adj, int* __restrict__ dropped) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= V) return;
    int s = shapeId[m];
    int v0 = polyStart[s], n = polyStart[s + 1] - v0;
    int a = m - v0;
    int pbase = packPoly[polyPack[s]];
    int part[MD - 2];
    int np = 0, drop = 0;
    auto insert = [&](int u) {
        if (np == MD - 2) {   // keep the MD-2 smallest ids
            if (u >= part[MD - 3]) { ++drop; return; }
            --np;
            ++drop;
        }
        int k = np;
        while (k > 0 && part[k - 1] > u) {
            part[k] = part[k - 1];
            --k;
        }
        part[k] = u;
        ++np;
    };
    int nb = min(bcnt[s], Kp);
    for (int r = 0; r < nb; ++r) {
        uint64_t k = bkey[(long long)s * Kp + r];
        if (key_i(k) == a) insert(polyStart[pbase + key_sj(k)] + key_j(k));
    }
    int nj = min(jcnt[s], Kp);
    for (int r = 0; r < nj; ++r) {
        uint64_t k = jkey[(long long)s * Kp + r];
        if (key_j(k) == a) insert(polyStart[pbase + key_si(k)] + key_i(k));
    }
    int* row = adj + (long long)m * MD;
    row[0] = v0 + (a + n - 1) % n;
    row[1] = v0 + (a + 1) % n;
    for (int k = 0; k < MD - 2; ++k) row[2 + k] = (k < np) ? part[k] : -1;
    if (drop) atomicAdd(dropped, drop);
}

// ---- input features --------------------------------------------------------------------
// x = (asinh(Fx / (ell^(e-1) s0)), asinh(Fy / (ell^(e-1) s1)), crossings(m) / s2,
//      asinh(E_poly / (n ell^e s3))),  ell = the packing's length unit, e = energyPow
__global__ void k_mpnn_features(int V, const int* __restrict__ shapeId, const int* __restrict__ polyStart,
                                const int* __restrict__ polyPack, const double2* __restrict__ F,
                                const double* __restrict__ epoly, const int* __restrict__ adj,
                                const double* __restrict__ ell, int energyPow, float s0, float s1, float s2,
                                float s3, float4* __restrict__ feat) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= V) return;
    int s = shapeId[m];
    double L = ell[polyPack[s]];
    double Lf = 1.0, Le;
    for (int k = 1; k < energyPow; ++k) Lf *= L;
    Le = Lf * L;
    int n = polyStart[s + 1] - polyStart[s];
    const int* row = adj + (long long)m * MD;
    int nc = 0;
    for (int k = 2; k < MD; ++k) nc += (row[k] >= 0);
    double2 f = F[m];
    float4 o;
    o.x = (float)asinh(f.x / (Lf * (double)s0));
    o.y = (float)asinh(f.y / (Lf * (double)s1));
    o.z = (float)nc / s2;
    o.w = (float)asinh(epoly[s] / ((double)n * Le * (double)s3));
    feat[m] = o;
}

// ---- forward ---------------------------------------------------------------------------
// h0 = relu(bf(x) bf(W0) + b0), thread per (vertex, channel), f = 0..3 in order
__global__ void k_mpnn_embed(int V, const float* __restrict__ prm, const float4* __restrict__ feat,
                             float* __restrict__ H0, float* __restrict__ Z0) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)V * MH) return;
    int m = (int)(t / MH), c = (int)(t % MH);
    const float* W0 = prm + PCP_MPNN_OFF_W0;
    float4 x = feat[m];
    float z = bfr(x.x) * bfr(W0[0 * MH + c]);
    z = z + bfr(x.y) * bfr(W0[1 * MH + c]);
    z = z + bfr(x.z) * bfr(W0[2 * MH + c]);
    z = z + bfr(x.w) * bfr(W0[3 * MH + c]);
    z = z + prm[PCP_MPNN_OFF_B0 + c];
    Z0[t] = z;
    H0[t] = fmaxf(z, 0.0f);
}

// One message-passing layer.  Wavefront w of the workgroup takes vertices
// [16 (4 b + w), +16); lane (r = lane % 16, g = lane / 16) owns columns 16 s + 4 g .. +3
// (s = 0..3) of vertex r for the A operand: its own h and the sum of its neighbours' h
// (slots in order).  X = bf16([h, agg]) is stored for the weight gradient.
__global__ __launch_bounds__(256) void k_mpnn_layer(int V, const float* __restrict__ W, const float* __restrict__ b,
                                                    const float* __restrict__ Hin, const int* __restrict__ adj,
                                                    float* __restrict__ Hout, float* __restrict__ Z,
                                                    unsigned short* __restrict__ X) {
    // W^T in LDS as bf16: sWt[n][k] (n = output column, k = input row), rows padded
    __shared__ unsigned short sWt[MH * (2 * MH + WPAD)];
    for (int e = threadIdx.x; e < 2 * MH * MH; e += blockDim.x) {
        int k = e / MH, n = e % MH;
        sWt[n * (2 * MH + WPAD) + k] = bf16_rne(W[e]);
    }
    __syncthreads();
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int r = lane & 15, g = lane >> 4;
    int row0 = (blockIdx.x * 4 + wave) * 16;
    if (row0 >= V) return;
    int row = row0 + r;
    bool ok = row < V;
    float4 h[4], a[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        h[s] = ok ? *(const float4*)(Hin + (long long)row * MH + 16 * s + 4 * g) : make_float4(0.f, 0.f, 0.f, 0.f);
        a[s] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    if (ok) {
        const int4* ar = (const int4*)(adj + (long long)row * MD);
        int nb[MD];
#pragma unroll
        for (int q = 0; q < MD / 4; ++q) {
            int4 t = ar[q];
            nb[4 * q] = t.x; nb[4 * q + 1] = t.y; nb[4 * q + 2] = t.z; nb[4 * q + 3] = t.w;
        }
#pragma unroll
        for (int d = 0; d < MD; ++d) {
            int u = nb[d];
            if (u >= 0) {
                const float* hu = Hin + (long long)u * MH + 4 * g;
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    float4 v = *(const float4*)(hu + 16 * s);
                    a[s].x = a[s].x + v.x; a[s].y = a[s].y + v.y; a[s].z = a[s].z + v.z; a[s].w = a[s].w + v.w;
                }
            }
        }
    }
    mpnn_s4 A[8];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        A[s] = pack_bf16(h[s].x, h[s].y, h[s].z, h[s].w);
        A[4 + s] = pack_bf16(a[s].x, a[s].y, a[s].z, a[s].w);
    }
    if (ok) {
        unsigned short* xr = X + (long long)row * (2 * MH) + 4 * g;
#pragma unroll
        for (int s = 0; s < 8; ++s) *(mpnn_s4*)(xr + 16 * s) = A[s];
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        mpnn_f4 acc = {0.f, 0.f, 0.f, 0.f};
        const unsigned short* wr = sWt + (16 * t + r) * (2 * MH + WPAD) + 4 * g;
#pragma unroll
        for (int s = 0; s < 8; s += 2) {
            mpnn_s8 B = cat_s4(*(const mpnn_s4*)(wr + 16 * s), *(const mpnn_s4*)(wr + 16 * s + 16));
            acc = mfma32(cat_s4(A[s], A[s + 1]), B, acc);
        }
        int col = 16 * t + r;
        float bias = b[col];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int ro = row0 + 4 * g + i;
            if (ro < V) {
                float z = acc[i] + bias;
                long long o = (long long)ro * MH + col;
                Z[o] = z;
                Hout[o] = Hin[o] + fmaxf(z, 0.0f);
            }
        }
    }
}

// y = bf(h3) bf(Wo) + bo, thread per vertex, channels in order; the workgroup's 128 rows of
// h3 are staged in LDS with coalesced loads (rows padded to 65 floats against bank conflicts)
__global__ __launch_bounds__(128) void k_mpnn_out(int V, const float* __restrict__ prm, const float* __restrict__ H3,
                                                  float2* __restrict__ out) {
    __shared__ float sh[128 * (MH + 1)];
    __shared__ float wo[2 * MH];
    int base = blockIdx.x * 128;
    for (int k = threadIdx.x; k < 128 * (MH / 4); k += 128) {
        int r = k / (MH / 4), q = k % (MH / 4);
        int m = base + r;
        float4 v = m < V ? *(const float4*)(H3 + (long long)m * MH + 4 * q) : make_float4(0.f, 0.f, 0.f, 0.f);
        float* d = sh + r * (MH + 1) + 4 * q;
        d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
    }
    wo[threadIdx.x] = bfr(prm[PCP_MPNN_OFF_WO + threadIdx.x]);
    __syncthreads();
    int m = base + threadIdx.x;
    if (m >= V) return;
    const float* h = sh + threadIdx.x * (MH + 1);
    float y0 = 0.0f, y1 = 0.0f;
    for (int c = 0; c < MH; ++c) {
        float hc = bfr(h[c]);
        y0 = y0 + hc * wo[2 * c];
        y1 = y1 + hc * wo[2 * c + 1];
    }
    out[m] = make_float2(y0 + prm[PCP_MPNN_OFF_BO], y1 + prm[PCP_MPNN_OFF_BO + 1]);
}

// ---- loss ------------------------------------------------------------------------------
// L = (1/V) sum_m |y_m - t_m|^2 ; dY = 2 (y - t) / V ; per-workgroup partial sums (double)
__global__ __launch_bounds__(256) void k_mpnn_loss(int V, const float2* __restrict__ y, const float2* __restrict__ tg,
                                                   float2* __restrict__ dy, double* __restrict__ part) {
    __shared__ double sh[256];
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (m < V) {
        float2 a = y[m], t = tg[m];
        float dx = a.x - t.x, dz = a.y - t.y;
        e = (double)dx * (double)dx + (double)dz * (double)dz;
        float sc = 2.0f / (float)V;
        dy[m] = make_float2(sc * dx, sc * dz);
    }
    sh[threadIdx.x] = e;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// grp[gi * width + e] = sum over the chunks q of group gi (MPNN_GRP consecutive chunks, in
// order) of part[q * width + e]; thread per (entry, group)
__global__ void k_mpnn_sum_groups(int nChunks, int width, const float* __restrict__ part, float* __restrict__ grp) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    int gi = blockIdx.y;
    if (e >= width) return;
    int q0 = gi * MPNN_GRP, q1 = min(nChunks, q0 + MPNN_GRP);
    float s = 0.0f;
    for (int q = q0; q < q1; ++q) s = s + part[(long long)q * width + e];
    grp[(long long)gi * width + e] = s;
}
// out[e] = sum over groups (in order) of grp[gi * width + e]
__global__ void k_mpnn_sum_chunks(int nGroups, int width, const float* __restrict__ grp, float* __restrict__ out) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= width) return;
    float s = 0.0f;
    for (int q = 0; q < nGroups; ++q) s = s + grp[(long long)q * width + e];
    out[e] = s;
}
// loss = scale * sum of the partials: lane t adds partials t, t + 256, ... in order, then a
// fixed pairwise tree over the 256 lane sums
__global__ __launch_bounds__(256) void k_mpnn_sum_loss(int n, const double* __restrict__ part, double scale,
                                                       double* __restrict__ out) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int q = threadIdx.x; q < n; q += 256) s = s + part[q];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0] * scale;
}

// ---- backward --------------------------------------------------------------------------
// dh3 = dY bf(Wo)^T (thread per (vertex, channel))
__global__ void k_mpnn_out_bwd(int V, const float* __restrict__ prm, const float2* __restrict__ dy,
                               float* __restrict__ dH) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)V * MH) return;
    int m = (int)(t / MH), c = (int)(t % MH);
    const float* Wo = prm + PCP_MPNN_OFF_WO;
    float2 g = dy[m];
    dH[t] = g.x * bfr(Wo[2 * c]) + g.y * bfr(Wo[2 * c + 1]);
}

// partial gradients of Wo (64x2) and bo over a chunk of MPNN_CHV vertices: entry e < 128 is
// Wo[e/2][e%2], e = 128, 129 is bo[e-128].  One wavefront per chunk, lane c owns channel c
// (and lane 0 the biases); each entry is summed over the chunk's vertices in order.
__global__ __launch_bounds__(64) void k_mpnn_out_grad(int V, const float* __restrict__ H3, const float2* __restrict__ dy,
                                                      float* __restrict__ part) {
    int c = threadIdx.x;
    int m0 = blockIdx.x * MPNN_CHV, m1 = min(V, m0 + MPNN_CHV);
    float s0 = 0.0f, s1 = 0.0f, b0 = 0.0f, b1 = 0.0f;
#pragma unroll 4
    for (int m = m0; m < m1; ++m) {
        float2 g = dy[m];
        float h = bfr(H3[(long long)m * MH + c]);
        s0 = s0 + h * g.x;
        s1 = s1 + h * g.y;
        b0 = b0 + g.x;
        b1 = b1 + g.y;
    }
    float* out = part + (long long)blockIdx.x * PCP_MPNN_OUT_NP;
    out[2 * c] = s0;
    out[2 * c + 1] = s1;
    if (c == 0) {
        out[2 * MH] = b0;
        out[2 * MH + 1] = b1;
    }
}

// Layer backward, input side.  dZ = dH ⊙ [Z > 0]; dX = bf(dZ) bf(W)^T (MFMA, K = 64);
// dHself = dH + dX[:, :64] (residual + own row), dA = dX[:, 64:] (to be gathered).
// Lane (r, g) holds dZ[row][16 s + 4 g .. +3], s = 0..3, as the A operand; the B operand
// B[k][j] = W[j][k] is read from W row-major in LDS.
__global__ __launch_bounds__(256) void k_mpnn_layer_bwd(int V, const float* __restrict__ W,
                                                        const float* __restrict__ dH, const float* __restrict__ Z,
                                                        float* __restrict__ dHself, float* __restrict__ dA) {
    __shared__ unsigned short sW[2 * MH * (MH + WPAD)];   // W[j][k], j < 128 inputs, k < 64 outputs
    for (int e = threadIdx.x; e < 2 * MH * MH; e += blockDim.x) {
        int j = e / MH, k = e % MH;
        sW[j * (MH + WPAD) + k] = bf16_rne(W[e]);
    }
    __syncthreads();
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int r = lane & 15, g = lane >> 4;
    int row0 = (blockIdx.x * 4 + wave) * 16;
    if (row0 >= V) return;
    int row = row0 + r;
    bool ok = row < V;
    mpnn_s4 A[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        float4 d = make_float4(0.f, 0.f, 0.f, 0.f), z = make_float4(0.f, 0.f, 0.f, 0.f);
        if (ok) {
            d = *(const float4*)(dH + (long long)row * MH + 16 * s + 4 * g);
            z = *(const float4*)(Z + (long long)row * MH + 16 * s + 4 * g);
        }
        A[s] = pack_bf16(z.x > 0.f ? d.x : 0.f, z.y > 0.f ? d.y : 0.f, z.z > 0.f ? d.z : 0.f,
                         z.w > 0.f ? d.w : 0.f);
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        mpnn_f4 acc = {0.f, 0.f, 0.f, 0.f};
        const unsigned short* wr = sW + (16 * t + r) * (MH + WPAD) + 4 * g;
#pragma unroll
        for (int s = 0; s < 4; s += 2) {
            mpnn_s8 B = cat_s4(*(const mpnn_s4*)(wr + 16 * s), *(const mpnn_s4*)(wr + 16 * s + 16));
            acc = mfma32(cat_s4(A[s], A[s + 1]), B, acc);
        }
        int col = 16 * t + r;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int ro = row0 + 4 * g + i;
            if (ro >= V) continue;
            if (t < 4) {
                long long o = (long long)ro * MH + col;
                dHself[o] = dH[o] + acc[i];
            } else {
                dA[(long long)ro * MH + (col - MH)] = acc[i];
            }
        }
    }
}

// dH_prev[m] = dHself[m] + sum over the slots of m (in order) of dA[u]: the transpose of
// the neighbour sum, as the adjacency is symmetric
__global__ void k_mpnn_gather_bwd(int V, const int* __restrict__ adj, const float* __restrict__ dHself,
                                  const float* __restrict__ dA, float* __restrict__ dHprev) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)V * MH) return;
    int m = (int)(t / MH), c = (int)(t % MH);
    const int4* row = (const int4*)(adj + (long long)m * MD);
    int nb[MD];
#pragma unroll
    for (int q = 0; q < MD / 4; ++q) {
        int4 r4 = row[q];
        nb[4 * q] = r4.x; nb[4 * q + 1] = r4.y; nb[4 * q + 2] = r4.z; nb[4 * q + 3] = r4.w;
    }
    float v[MD];
#pragma unroll
    for (int d = 0; d < MD; ++d) v[d] = dA[(long long)(nb[d] >= 0 ? nb[d] : m) * MH + c];   // all loads in flight
    float s = 0.0f;
#pragma unroll
    for (int d = 0; d < MD; ++d) s = s + (nb[d] >= 0 ? v[d] : 0.0f);   // empty slots add +0.0
    dHprev[t] = dHself[t] + s;
}

// Weight gradient of a message layer over one chunk of MPNN_CH vertices:
// dW[k][n] = sum_m X[m][k] bf(dZ[m][n]) (MFMA, M = 128 inputs, N = 64 outputs, K = vertices)
// and db[n] = sum_m dZ[m][n] (float32, rows in order).  32 vertices at a time are staged
// in LDS (one 32-deep MFMA step: lane group g takes vertices 4g..4g+3 and 16+4g..16+4g+3 of
// the block); wavefront w accumulates the input tiles 2w, 2w+1 against the 4 output tiles.
// part row: 128*64 entries of dW (row-major) then 64 of db.
__global__ __launch_bounds__(256) void k_mpnn_layer_grad(int V, const unsigned short* __restrict__ X,
                                                         const float* __restrict__ dH, const float* __restrict__ Z,
                                                         float* __restrict__ part) {
    __shared__ unsigned short sX[32 * (2 * MH + WPAD)];
    __shared__ unsigned short sD[32 * (MH + WPAD)];
    __shared__ float sB[32 * MH];
    int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    int r = lane & 15, g = lane >> 4;
    int m0 = blockIdx.x * MPNN_CH, m1 = min(V, m0 + MPNN_CH);
    mpnn_f4 acc[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[a][t] = (mpnn_f4){0.f, 0.f, 0.f, 0.f};
    // bias partial: thread tid owns column tid % 64 of the rows with (row % 16) / 4 == tid / 64,
    // added in row order
    float bsum = 0.0f;
    for (int base = m0; base < m1; base += 32) {
        // stage X rows (32 x 128 bf16): thread -> row tid / 8, 16 elements at 16 (tid % 8)
        {
            int rr = tid >> 3, c16 = (tid & 7) * 16;
            int m = base + rr;
            uint4 v0 = make_uint4(0u, 0u, 0u, 0u), v1 = make_uint4(0u, 0u, 0u, 0u);
            if (m < m1) {
                v0 = *(const uint4*)(X + (long long)m * (2 * MH) + c16);
                v1 = *(const uint4*)(X + (long long)m * (2 * MH) + c16 + 8);
            }
            unsigned short* d = sX + rr * (2 * MH + WPAD) + c16;
            *(uint2*)d = make_uint2(v0.x, v0.y);
            *(uint2*)(d + 4) = make_uint2(v0.z, v0.w);
            *(uint2*)(d + 8) = make_uint2(v1.x, v1.y);
            *(uint2*)(d + 12) = make_uint2(v1.z, v1.w);
        }
        // stage dZ rows (32 x 64): thread -> row tid / 8, 8 columns at 8 (tid % 8)
        {
            int rr = tid >> 3, c8 = (tid & 7) * 8;
            int m = base + rr;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                float4 d = make_float4(0.f, 0.f, 0.f, 0.f), z = make_float4(0.f, 0.f, 0.f, 0.f);
                if (m < m1) {
                    d = *(const float4*)(dH + (long long)m * MH + c8 + 4 * h);
                    z = *(const float4*)(Z + (long long)m * MH + c8 + 4 * h);
                }
                float4 dz = make_float4(z.x > 0.f ? d.x : 0.f, z.y > 0.f ? d.y : 0.f, z.z > 0.f ? d.z : 0.f,
                                        z.w > 0.f ? d.w : 0.f);
                *(mpnn_s4*)(sD + rr * (MH + WPAD) + c8 + 4 * h) = pack_bf16(dz.x, dz.y, dz.z, dz.w);
                *(float4*)(sB + rr * MH + c8 + 4 * h) = dz;
            }
        }
        __syncthreads();
        {
            // thread tid: column tid % 64, rows 4 (tid / 64) .. +3 of each half of this block of 32
            int c = tid & 63, q = tid >> 6;
#pragma unroll
            for (int i = 0; i < 4; ++i) bsum = bsum + sB[(4 * q + i) * MH + c];
#pragma unroll
            for (int i = 0; i < 4; ++i) bsum = bsum + sB[(16 + 4 * q + i) * MH + c];
        }
        mpnn_s8 Bop[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            mpnn_s8 b;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                b[e] = (short)sD[(4 * g + e) * (MH + WPAD) + 16 * t + r];
                b[4 + e] = (short)sD[(16 + 4 * g + e) * (MH + WPAD) + 16 * t + r];
            }
            Bop[t] = b;
        }
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            int kcol = 16 * (2 * wave + a) + r;
            mpnn_s8 Aop;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                Aop[e] = (short)sX[(4 * g + e) * (2 * MH + WPAD) + kcol];
                Aop[4 + e] = (short)sX[(16 + 4 * g + e) * (2 * MH + WPAD) + kcol];
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[a][t] = mfma32(Aop, Bop[t], acc[a][t]);
        }
        __syncthreads();
    }
    float* out = part + (long long)blockIdx.x * PCP_MPNN_LAYER_NP;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                int k = 16 * (2 * wave + a) + 4 * g + i, n = 16 * t + r;
                out[k * MH + n] = acc[a][t][i];
            }
    // bias: the four row groups of each column, added in order
    sB[tid] = bsum;
    __syncthreads();
    if (tid < MH) out[2 * MH * MH + tid] = ((sB[tid] + sB[MH + tid]) + sB[2 * MH + tid]) + sB[3 * MH + tid];
}

// Embedding gradient over one chunk of MPNN_CHV vertices: dZ0 = dH0 ⊙ [Z0 > 0]; entry e < 256 is
// W0[e/64][e%64] = sum_m bf(x_m[e/64]) dZ0[m][e%64], e = 256..319 is b0[e-256].  One wavefront
// per chunk, lane c owns the five entries of channel c; vertices in order.
__global__ __launch_bounds__(64) void k_mpnn_embed_grad(int V, const float4* __restrict__ feat,
                                                        const float* __restrict__ dH0, const float* __restrict__ Z0,
                                                        float* __restrict__ part) {
    int c = threadIdx.x;
    int m0 = blockIdx.x * MPNN_CHV, m1 = min(V, m0 + MPNN_CHV);
    float w0 = 0.0f, w1 = 0.0f, w2 = 0.0f, w3 = 0.0f, bs = 0.0f;
#pragma unroll 4
    for (int m = m0; m < m1; ++m) {
        long long o = (long long)m * MH + c;
        float dz = Z0[o] > 0.f ? dH0[o] : 0.f;
        float4 x = feat[m];
        w0 = w0 + bfr(x.x) * dz;
        w1 = w1 + bfr(x.y) * dz;
        w2 = w2 + bfr(x.z) * dz;
        w3 = w3 + bfr(x.w) * dz;
        bs = bs + dz;
    }
    float* out = part + (long long)blockIdx.x * PCP_MPNN_EMBED_NP;
    out[0 * MH + c] = w0;
    out[1 * MH + c] = w1;
    out[2 * MH + c] = w2;
    out[3 * MH + c] = w3;
    out[4 * MH + c] = bs;
}

// ---- optimizer, targets, application -----------------------------------------------------
// Adam (Kingma & Ba 2015) with bias correction, float32
__global__ void k_mpnn_adam(int n, float* __restrict__ p, const float* __restrict__ gr, float* __restrict__ mo,
                            float* __restrict__ ve, float lr, float b1, float b2, float eps, float c1, float c2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float g = gr[i];
    float m = b1 * mo[i] + (1.0f - b1) * g;
    float v = b2 * ve[i] + (1.0f - b2) * (g * g);
    mo[i] = m;
    ve[i] = v;
    float mh = m / c1, vh = v / c2;
    p[i] = p[i] - lr * mh / (sqrtf(vh) + eps);
}

// target_m = minimum image (xTo_m - xFrom_m) / (ell * outScale)
__global__ void k_mpnn_targets(int V, const int* __restrict__ shapeId, const int* __restrict__ polyPack,
                               const double2* __restrict__ xFrom, const double2* __restrict__ xTo,
                               const double* __restrict__ ell, double outScale, float2* __restrict__ tg) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= V) return;
    double L = ell[polyPack[shapeId[m]]] * outScale;
    double2 a = xFrom[m], b = xTo[m];
    double dx = b.x - a.x, dy = b.y - a.y;
    dx = dx - rint(dx);
    dy = dy - rint(dy);
    tg[m] = make_float2((float)(dx / L), (float)(dy / L));
}

// x_m <- frac(x_m + ell outScale y_m) for vertices of packings with active != 0
__global__ void k_mpnn_apply(int V, const int* __restrict__ shapeId, const int* __restrict__ polyPack,
                             const int* __restrict__ active, const float2* __restrict__ y,
                             const double* __restrict__ ell, double outScale, double2* __restrict__ x) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= V) return;
    int p = polyPack[shapeId[m]];
    if (active && !active[p]) return;
    double L = ell[p] * outScale;
    float2 d = y[m];
    double2 xi = x[m];
    double px = xi.x + L * (double)d.x, py = xi.y + L * (double)d.y;
    px = px - floor(px);
    py = py - floor(py);
    if (px >= 1.0) px = 0.0;
    if (py >= 1.0) py = 0.0;
    x[m] = make_double2(px, py);
}

// Guard of the warm start.  An edge-crossing energy is blind to a polygon that lies inside
// another (no edge crosses), and while one vertex still pokes out the whole polygon's area
// counts, so FIRE pulls a polygon that is mostly inside another all the way in and cannot
// bring it out.  Where the prediction puts one polygon's bounding circle inside another's,
// the two polygons take back their positions; where that leaves such a pair in the packing
// (with a neighbour that moved), the whole packing takes back its positions.
// Thread per polygon A: the polygons of larger id in the 3x3 block of polygon cells around
// A (k_poly_cells of the current positions; side >= the largest circle diameter, so
// circles that meet lie in neighbouring cells); a pair is contained when
// |c_A - c_B| + min(r_A, r_B) < kappa * max(r_A, r_B) (minimum image).  PER_POLY: both
// polygons of every contained pair get polyFlag = 1 and their packing flag = 1; else the
// packing gets flag = 2.  Polygons that wind around the torus are not judged.
template <bool PER_POLY>
__global__ void k_contained(int N, const int* __restrict__ polyPack, const int* __restrict__ packPoly,
                            const double4* __restrict__ bounds, const int* __restrict__ pcStart,
                            const int* __restrict__ pcSorted, const int* __restrict__ packB, double kappa, int* __restrict__ flag,
                            int* __restrict__ polyFlag) {
    int A = blockIdx.x * blockDim.x + threadIdx.x;
    if (A >= N) return;
    int pk = polyPack[A];
    int s0 = packPoly[pk];
    int Bp = packB[pk] & 0xFFFF;
    double4 bA = bounds[A];
    if (bA.w != 0.0) return;
    long long gbase = (long long)pk * (PC_CELLS + 1);
    int ax = pc_coord(bA.x, Bp), ay = pc_coord(bA.y, Bp);
    int ncell = (Bp == 1) ? 1 : 9;
    for (int k = 0; k < ncell; ++k) {
        int cx = ax, cy = ay;
        if (Bp > 1) {
            cx = ax + k / 3 - 1; cx += (cx < 0) ? Bp : 0; cx -= (cx >= Bp) ? Bp : 0;
            cy = ay + k % 3 - 1; cy += (cy < 0) ? Bp : 0; cy -= (cy >= Bp) ? Bp : 0;
        }
        int cell = cy * Bp + cx;
        int q1 = pcStart[gbase + cell + 1];
        for (int q = pcStart[gbase + cell]; q < q1; ++q) {
            int Bq = pcSorted[s0 + q];
            if (Bq <= A) continue;
            double4 b = bounds[Bq];
            if (b.w != 0.0) continue;
            double lim = kappa * fmax(bA.z, b.z) - fmin(bA.z, b.z);
            if (!(lim > 0.0)) continue;
            double dx = wrap(b.x - bA.x), dy = wrap(b.y - bA.y);
            if (dx * dx + dy * dy < lim * lim) {
                if (!PER_POLY) {
                    flag[pk] = 2;
                    return;
                }
                polyFlag[A] = 1;
                polyFlag[Bq] = 1;
                flag[pk] = 1;
            }
        }
    }
}

// x = xPrev for the vertices of flagged polygons (PER_POLY) or packings (flag == 2)
template <bool PER_POLY>
__global__ void k_take_back(int V, const int* __restrict__ shapeId, const int* __restrict__ polyPack,
                            const int* __restrict__ flag, const int* __restrict__ polyFlag,
                            const double2* __restrict__ xPrev, double2* __restrict__ x) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= V) return;
    int s = shapeId[m];
    if (PER_POLY ? polyFlag[s] != 0 : flag[polyPack[s]] == 2) x[m] = xPrev[m];
}

// work buffer slices (floats): H0..H3, Z0..Z3 (V x 64 each), X1..X3 (V x 128 bf16 each)
struct MpnnWork {
    float* H[4];
    float* Z[4];
    unsigned short* X[4];
};
static MpnnWork mpnn_slices(float* work, int V) {
    MpnnWork w;
    long long S = (long long)V * MH;
    for (int l = 0; l < 4; ++l) {
        w.H[l] = work + l * S;
        w.Z[l] = work + (4 + l) * S;
    }
    w.X[0] = nullptr;
    for (int l = 1; l < 4; ++l) w.X[l] = (unsigned short*)(work + (8 + (l - 1)) * S);
    return w;
}
static int mpnn_chunks(int V) { return (V + MPNN_CH - 1) / MPNN_CH; }
static int mpnn_chunks_v(int V) { return (V + MPNN_CHV - 1) / MPNN_CHV; }
static int mpnn_groups(int nChunks) { return (nChunks + MPNN_GRP - 1) / MPNN_GRP; }
static long long mpnn_part_floats(int V) {
    return std::max((long long)mpnn_chunks(V) * PCP_MPNN_LAYER_NP, (long long)mpnn_chunks_v(V) * PCP_MPNN_EMBED_NP);
}
static long long mpnn_group_floats(int V) {
    return std::max((long long)mpnn_groups(mpnn_chunks(V)) * PCP_MPNN_LAYER_NP,
                    (long long)mpnn_groups(mpnn_chunks_v(V)) * PCP_MPNN_EMBED_NP);
}
// out[0..width) = sum of the chunk partials part[nChunks][width] (groups, then groups in order)
static int mpnn_sum_parts(pcp_ctx* c, int nChunks, int width, const float* part, float* grp, float* out) {
    int ng = mpnn_groups(nChunks);
    PCP_LAUNCH(c, k_mpnn_sum_groups, dim3(nblk(width), ng), dim3(BLK), nChunks, width, part, grp);
    PCP_LAUNCH(c, k_mpnn_sum_chunks, dim3(nblk(width)), dim3(BLK), ng, width, grp, out);
    return PCP_OK;
}

}  // namespace pcp

extern "C" {

int pcp_mpnn_sizes(int V, long long* workFloats, long long* gradWorkFloats) {
    if (V < 0) {
        pcp::set_error("pcp_mpnn_sizes: V < 0");
        return PCP_ERR_ARG;
    }
    long long S = (long long)V * PCP_MPNN_H;
    if (workFloats) *workFloats = 11 * S;
    // dH (two buffers), dHself, dA (V x 64 each), loss partials (V/256 + 2 doubles),
    // weight-gradient partials (one row per chunk), their group sums
    if (gradWorkFloats)
        *gradWorkFloats = 4 * S + 2LL * (V / 256 + 2) + pcp::mpnn_part_floats(V) + pcp::mpnn_group_floats(V);
    return PCP_OK;
}

int pcp_mpnn_graph(pcp_ctx* c, int* adj, int* dropped) {
    using namespace pcp;
    if (c->V == 0) {
        if (dropped) *dropped = 0;
        return PCP_OK;
    }
    if (!c->dBcnt || !c->dJcnt) {
        set_error("pcp_mpnn_graph: no records (evaluate forces first)");
        return PCP_ERR_ARG;
    }
    // the counter goes into the last scalar slot of the context
    int* dDrop = c->dScal + SC_MPNN_DROP;
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, dDrop, 1LL, 0);
    PCP_LAUNCHED();
    PCP_LAUNCH(c, k_mpnn_graph, dim3(nblk(c->V)), dim3(BLK), c->V, c->Kp, c->dPolyStart, c->dShapeId, c->dPolyPack,
               c->dPackPoly, c->dBcnt, c->dBKey, c->dJcnt, c->dJKey, adj, dDrop);
    if (dropped) {
        PCP_HIP(hipMemcpyAsync(dropped, dDrop, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        PCP_HIP(hipStreamSynchronize(c->stream));
    }
    return PCP_OK;
}

int pcp_mpnn_features(pcp_ctx* c, const pcp_state* st, const int* adj, const double* ell, int energyPow,
                      const float* inScale, float* feat) {
    using namespace pcp;
    if (energyPow < 1 || energyPow > 8 || !inScale) {
        set_error("pcp_mpnn_features: energyPow must be in 1..8 and inScale given");
        return PCP_ERR_ARG;
    }
    if (c->V == 0) return PCP_OK;
    PCP_LAUNCH(c, k_mpnn_features, dim3(nblk(c->V)), dim3(BLK), c->V, c->dShapeId, c->dPolyStart, c->dPolyPack,
               (const double2*)st->force, c->dEpoly, adj, ell, energyPow, inScale[0], inScale[1], inScale[2],
               inScale[3], (float4*)feat);
    return PCP_OK;
}

int pcp_mpnn_forward(pcp_ctx* c, int V, const float* params, const float* feat, const int* adj, float* work,
                     float* out) {
    using namespace pcp;
    if (V < 0) {
        set_error("pcp_mpnn_forward: V < 0");
        return PCP_ERR_ARG;
    }
    if (V == 0) return PCP_OK;
    MpnnWork w = mpnn_slices(work, V);
    long long S = (long long)V * MH;
    PCP_LAUNCH(c, k_mpnn_embed, dim3(nblk(S)), dim3(BLK), V, params, (const float4*)feat, w.H[0], w.Z[0]);
    int G = (V + 63) / 64;
    for (int l = 1; l <= 3; ++l) {
        const float* W = params + PCP_MPNN_OFF_W1 + (l - 1) * PCP_MPNN_LAYER_NP;
        PCP_LAUNCH(c, k_mpnn_layer, dim3(G), dim3(256), V, W, W + 2 * MH * MH, w.H[l - 1], adj, w.H[l], w.Z[l],
                   w.X[l]);
    }
    PCP_LAUNCH(c, k_mpnn_out, dim3((V + 127) / 128), dim3(128), V, params, w.H[3], (float2*)out);
    return PCP_OK;
}

int pcp_mpnn_loss(pcp_ctx* c, int V, const float* out, const float* target, float* dOut, float* gradWork,
                  double* loss) {
    using namespace pcp;
    if (V <= 0) {
        set_error("pcp_mpnn_loss: V must be positive");
        return PCP_ERR_ARG;
    }
    long long S = (long long)V * MH;
    double* part = (double*)(gradWork + 4 * S);
    int nb = nblk(V);
    PCP_LAUNCH(c, k_mpnn_loss, dim3(nb), dim3(256), V, (const float2*)out, (const float2*)target, (float2*)dOut,
               part);
    double* dl = part + nb;
    PCP_LAUNCH(c, k_mpnn_sum_loss, dim3(1), dim3(256), nb, part, 1.0 / (double)V, dl);
    if (loss) {
        PCP_HIP(hipMemcpyAsync(loss, dl, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        PCP_HIP(hipStreamSynchronize(c->stream));
    }
    return PCP_OK;
}

int pcp_mpnn_backward(pcp_ctx* c, int V, const float* params, const float* feat, const int* adj, const float* work,
                      const float* dOut, float* gradWork, float* grad) {
    using namespace pcp;
    if (V <= 0) {
        set_error("pcp_mpnn_backward: V must be positive");
        return PCP_ERR_ARG;
    }
    MpnnWork w = mpnn_slices((float*)work, V);
    long long S = (long long)V * MH;
    float* dHa = gradWork;
    float* dHb = gradWork + S;
    float* dHs = gradWork + 2 * S;
    float* dA = gradWork + 3 * S;
    float* part = gradWork + 4 * S + 2LL * (V / 256 + 2);
    float* grp = part + mpnn_part_floats(V);
    int nc = mpnn_chunks(V), ncv = mpnn_chunks_v(V);
    int G = (V + 63) / 64;
    // output layer
    PCP_LAUNCH(c, k_mpnn_out_grad, dim3(ncv), dim3(64), V, w.H[3], (const float2*)dOut, part);
    PCP_TRY(mpnn_sum_parts(c, ncv, PCP_MPNN_OUT_NP, part, grp, grad + PCP_MPNN_OFF_WO));
    PCP_LAUNCH(c, k_mpnn_out_bwd, dim3(nblk(S)), dim3(BLK), V, params, (const float2*)dOut, dHa);
    float* dH = dHa;
    float* dHn = dHb;
    for (int l = 3; l >= 1; --l) {
        const float* W = params + PCP_MPNN_OFF_W1 + (l - 1) * PCP_MPNN_LAYER_NP;
        PCP_LAUNCH(c, k_mpnn_layer_grad, dim3(nc), dim3(256), V, w.X[l], dH, w.Z[l], part);
        PCP_TRY(mpnn_sum_parts(c, nc, PCP_MPNN_LAYER_NP, part, grp, grad + PCP_MPNN_OFF_W1 + (l - 1) * PCP_MPNN_LAYER_NP));
        PCP_LAUNCH(c, k_mpnn_layer_bwd, dim3(G), dim3(256), V, W, dH, w.Z[l], dHs, dA);
        PCP_LAUNCH(c, k_mpnn_gather_bwd, dim3(nblk(S)), dim3(BLK), V, adj, dHs, dA, dHn);
        float* t = dH;
        dH = dHn;
        dHn = t;
    }
    PCP_LAUNCH(c, k_mpnn_embed_grad, dim3(ncv), dim3(64), V, (const float4*)feat, dH, w.Z[0], part);
    PCP_TRY(mpnn_sum_parts(c, ncv, PCP_MPNN_EMBED_NP, part, grp, grad + PCP_MPNN_OFF_W0));
    return PCP_OK;
}

int pcp_mpnn_adam(pcp_ctx* c, float* params, const float* grad, float* m, float* v, int step, float lr,
                  float beta1, float beta2, float eps) {
    using namespace pcp;
    if (step < 1) {
        set_error("pcp_mpnn_adam: step counts from 1");
        return PCP_ERR_ARG;
    }
    float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    PCP_LAUNCH(c, k_mpnn_adam, dim3(nblk(PCP_MPNN_NPARAM)), dim3(BLK), PCP_MPNN_NPARAM, params, grad, m, v, lr, beta1,
               beta2, eps, c1, c2);
    return PCP_OK;
}

int pcp_mpnn_targets(pcp_ctx* c, const double* xFrom, const double* xTo, const double* ell, double outScale,
                     float* target) {
    using namespace pcp;
    if (c->V == 0) return PCP_OK;
    PCP_LAUNCH(c, k_mpnn_targets, dim3(nblk(c->V)), dim3(BLK), c->V, c->dShapeId, c->dPolyPack,
               (const double2*)xFrom, (const double2*)xTo, ell, outScale, (float2*)target);
    return PCP_OK;
}

int pcp_mpnn_apply(pcp_ctx* c, double* pos, const float* out, const double* ell, double outScale,
                   const int* active) {
    using namespace pcp;
    if (c->V == 0) return PCP_OK;
    PCP_LAUNCH(c, k_mpnn_apply, dim3(nblk(c->V)), dim3(BLK), c->V, c->dShapeId, c->dPolyPack, active,
               (const float2*)out, ell, outScale, (double2*)pos);
    return PCP_OK;
}

int pcp_mpnn_guard(pcp_ctx* c, double* pos, const double* xPrev, double kappa, int* flags, int* nTaken) {
    using namespace pcp;
    if (!(kappa >= 0.0 && kappa <= 2.0)) return fail_arg("pcp_mpnn_guard: kappa must lie in [0, 2]");
    if (!flags) return fail_arg("pcp_mpnn_guard: flags is NULL");
    if (c->V > 0) {
        if (!c->dGuard) PCP_TRY(dalloc(&c->dGuard, std::max(c->N, 1)));
        PCP_LAUNCH(c, k_fill_int, dim3(nblk(c->P)), dim3(BLK), flags, (long long)c->P, 0);
        PCP_LAUNCH(c, k_fill_int, dim3(nblk(c->N)), dim3(BLK), c->dGuard, (long long)c->N, 0);
        // round 1 the contained pairs' polygons, round 2 the packings still holding one; each
        // on the bounding circles and polygon cells of the current positions (the first two
        // kernels of a detection; the records they reset are stale after a move anyway)
        for (int round = 0; round < 2; ++round) {
            PCP_TRY(launch_bounds(c, (const double2*)pos, nullptr));
            PCP_LAUNCH(c, k_poly_cells, dim3(c->P), dim3(PP_BLK), c->dPackPoly, c->dBounds, c->dPcStart,
                       c->dPcSorted, c->dPackB, (const int*)nullptr, (int*)nullptr, (int*)nullptr);
            if (round == 0) {
                PCP_LAUNCH(c, k_contained<true>, dim3(nblk(c->N)), dim3(BLK), c->N, c->dPolyPack, c->dPackPoly,
                           c->dBounds, c->dPcStart, c->dPcSorted, c->dPackB, kappa, flags, c->dGuard);
                PCP_LAUNCH(c, k_take_back<true>, dim3(nblk(c->V)), dim3(BLK), c->V, c->dShapeId, c->dPolyPack,
                           flags, c->dGuard, (const double2*)xPrev, (double2*)pos);
            } else {
                PCP_LAUNCH(c, k_contained<false>, dim3(nblk(c->N)), dim3(BLK), c->N, c->dPolyPack, c->dPackPoly,
                           c->dBounds, c->dPcStart, c->dPcSorted, c->dPackB, kappa, flags, c->dGuard);
                PCP_LAUNCH(c, k_take_back<false>, dim3(nblk(c->V)), dim3(BLK), c->V, c->dShapeId, c->dPolyPack,
                           flags, c->dGuard, (const double2*)xPrev, (double2*)pos);
            }
        }
    }
    if (nTaken) {
        std::vector<int> h(std::max(c->P, 1), 0);
        if (c->V > 0) {
            PCP_HIP(hipMemcpyAsync(h.data(), flags, sizeof(int) * c->P, hipMemcpyDeviceToHost, c->stream));
            PCP_HIP(hipStreamSynchronize(c->stream));
        }
        int n = 0;
        for (int p = 0; p < c->P; ++p) n += h[p] ? 1 : 0;
        *nTaken = n;
    }
    return PCP_OK;
}

}  // extern "C"
