// One Gauss-Newton iteration on the device: term residuals / Jacobians,
// J^T J diagonal + J^T r + block-Jacobi inverse per node, J^T J off-diagonal
// blocks per node pair, block-sparse PCG and the warp-field update.
#include "common.h"
#include "device_util.h"
#include "eval.h"
#include "terms.h"
#include "tuning.h"

namespace mvdr {

// One thread per global term; smooth terms forward-warp the graph edge
// (NodeGraphSmoothHandler.cu:10-43), feature / cross terms their canonical
// vertex (SparseCorrespondenceHandler.cu:193-220).
template <typename R>
__global__ void __launch_bounds__(256) eval_terms_kernel(const Buffers* __restrict__ bs, uint32_t which) {
    const Buffers b = frame_buffers(bs);
    const DQ* __restrict__ se3 = which ? b.se3[1] : b.se3[0];
    const FrameDev& f = *b.frame;
    const Counts& c = *b.counts;
    if (c.error) return;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < c.off[4]; t += stride) {
        if (t < c.off[0]) {
            const ushort4 px = b.dense_px[t];
            ScalarRec rec = b.dense[t];
            eval_dense<R>(f, se3, px, rec);
            b.dense[t].r = rec.r;
#pragma unroll
            for (int k = 0; k < 6; ++k) b.dense[t].J[k] = rec.J[k];
            if (b.density) {
                ScalarRec d = b.density[t];
                eval_image<R>(f, se3, px, true, d);
                b.density[t].r = d.r;
#pragma unroll
                for (int k = 0; k < 6; ++k) b.density[t].J[k] = d.J[k];
            }
        } else if (t < c.off[1]) {
            const uint32_t e = t - c.off[0];
            b.smooth[e] = eval_smooth<R>(f, se3, e);
        } else if (t < c.off[2]) {
            const uint32_t k = t - c.off[1];
            ScalarRec rec = b.fg[k];
            eval_image<R>(f, se3, b.fg_px[k], false, rec);
            b.fg[k].r = rec.r;
#pragma unroll
            for (int q = 0; q < 6; ++q) b.fg[k].J[q] = rec.J[q];
        } else {
            const bool is_cross = t >= c.off[3];
            const uint32_t k = is_cross ? t - c.off[3] : t - c.off[2];
            P2PRec* recs = is_cross ? b.cross : b.feat;
            const float4 can = is_cross ? b.cross_canon[k] : b.feat_canon[k];
            recs[k].warped = eval_warped<R>(se3, recs[k].knn, recs[k].w, can);
        }
    }
}

// -------------------------------------------------- 6x6 SPD inverse (Cholesky)
// inv(A) = L^-T L^-1 with A = L L^T; A is symmetric positive definite (the
// reference adds 1e-2 to every diagonal, PreconditionerBlocked.cu:272-275).
// In FP64 the six pivots are inverted once and the other divisions become
// products; the FP32 instance divides as the reference does.
template <typename T>
__device__ inline void spd_inverse6(const float* A, float* out, double* out64, double* linv64 = nullptr) {
    constexpr bool kRecip = sizeof(T) == 8;
    T L[21];  // packed lower triangle, row-major
    T rd[6];  // 1 / L(r, r) (FP64)
#define LI(r, c) L[(r) * ((r) + 1) / 2 + (c)]
#pragma unroll
    for (int r = 0; r < 6; ++r) {
#pragma unroll
        for (int col = 0; col <= r; ++col) {
            T s = (T)A[6 * r + col];
#pragma unroll
            for (int k = 0; k < col; ++k) s -= LI(r, k) * LI(col, k);
            if (r == col) {
                LI(r, r) = s > (T)0 ? sqrt(s) : (T)1e-30;
                if (kRecip) rd[r] = (T)1 / LI(r, r);
            } else {
                LI(r, col) = kRecip ? s * rd[col] : s / LI(col, col);
            }
        }
    }
    // X = L^-1 (lower), in place of a dense 6x6
    T X[36];
#pragma unroll
    for (int c2 = 0; c2 < 6; ++c2)
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            T s = (r == c2) ? (T)1 : (T)0;
#pragma unroll
            for (int k = 0; k < r; ++k) s -= LI(r, k) * X[6 * k + c2];
            X[6 * r + c2] = kRecip ? s * rd[r] : s / LI(r, r);
        }
#undef LI
    if (linv64)
        for (int a = 0; a < 6; ++a)
            for (int k = 0; k <= a; ++k) linv64[tri_index(a, k)] = (double)X[6 * a + k];
    if (!out && !out64) return;
    // inv = X^T X
    for (int r = 0; r < 6; ++r)
        for (int col = 0; col < 6; ++col) {
            T s = 0;
            for (int k = 0; k < 6; ++k) s += X[6 * k + r] * X[6 * k + col];
            if (out) out[6 * r + col] = (float)s;
            if (out64) out64[6 * r + col] = (double)s;
        }
}

// ------------------------------------------------- J^T J, J^T r assembly
// Segmented reductions over the flat index arrays, one wave per 64 entries:
//   Node2Term entries (node-major)     -> D_n = sum lambda^2 J^T J and
//     b_n = -sum lambda^2 J^T r (computeBlockDiagonalPreconditionerKernel,
//     PreconditionerBlocked.cu:172-284; computeJtResidualWithIndexKernel,
//     JtDotResidualIndexed.cu:116-224),
//   NodePair2Term entries (pair-major) -> B_ij = sum lambda^2 J_i^T J_j
//     (computeJtJNonDiagonalBlockKernel, JtJMaterializer.cu:270-365).
// Every lane evaluates its entry's contribution in node-centred coordinates;
// an inclusive segmented scan across the wave (runs of equal node / pair)
// leaves each run's sum in its last lane.  A run that lies inside the chunk is
// final and written in place; a run cut by a chunk boundary leaves a partial
// (slot A: the run holding lane 0, slot B: the run holding lane 63) that
// assemble_fixup_kernel adds up in chunk order.  Work per wave is fixed
// whatever the row and pair lengths; summation order is fixed.
constexpr uint32_t kDiagVals = 27;  // 21 upper-triangle entries + 6 of J^T r
constexpr uint32_t kPairVals = 36;

// Run sums of a chunk: lane e holds the contribution v of entry e (key: its
// node or pair; runs of a key are contiguous).  The wave stores v in LDS, then
// lane q adds component q down each run in entry order -- a fixed order -- and
// emit(key, first, last, sum) is called once per run by every lane (lanes
// q >= V carry no sum).  Column reads use an odd row stride: no bank conflicts.
constexpr uint32_t kRedStride = 37;

template <int V, typename Emit>
__device__ inline void chunk_runs(const float (&v)[V], uint32_t key, bool valid, uint32_t lane, float* lds, Emit emit) {
#pragma unroll
    for (int q = 0; q < V; ++q) lds[lane * kRedStride + q] = v[q];
    const uint32_t key_next = (uint32_t)__shfl_down((int)key, 1, 64);
    uint64_t ends = __ballot(valid && (lane == 63 || key_next != key));
    wave_lds_sync();
    uint32_t first = 0;
    while (ends) {
        const uint32_t last = (uint32_t)__builtin_ctzll(ends);
        ends &= ends - 1;
        const uint32_t rkey = (uint32_t)__shfl((int)key, (int)last, 64);
        float acc = 0.f;
        if (lane < (uint32_t)V) {
            // eight reads in flight, added in entry order
            uint32_t e = first;
            for (; e + 8 <= last + 1; e += 8) {
                float x[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) x[u] = lds[(e + u) * kRedStride + lane];
#pragma unroll
                for (int u = 0; u < 8; ++u) acc += x[u];
            }
            for (; e <= last; ++e) acc += lds[e * kRedStride + lane];
        }
        emit(rkey, first, last, acc);
        first = last + 1;
    }
    wave_lds_sync();  // the wave's LDS rows are rewritten by its next chunk
}

// Position (a, k), a <= k, of the q-th entry of a row-major upper triangle of 6x6.
__device__ inline void triu_index(uint32_t q, uint32_t& a, uint32_t& k) {
    a = 0;
    uint32_t rest = q;
    while (rest >= 6 - a) {
        rest -= 6 - a;
        ++a;
    }
    k = a + rest;
}

__device__ inline void diag_chunk(const Buffers& b, const Counts& c, const Penalty& pen, uint32_t ch, uint32_t lane,
                                  float* lds) {
    const FrameDev& f = *b.frame;
    const uint32_t total = c.kv1;
    const uint32_t e = ch * 64 + lane;
    const bool valid = e < total;
    const uint32_t key = valid ? b.n2t_node[e] : 0xFFFFFFFFu;
    const uint32_t prev_key = ch > 0 ? b.n2t_node[ch * 64 - 1] : 0xFFFFFFFFu;
    const uint32_t next_key = ch * 64 + 64 < total ? b.n2t_node[ch * 64 + 64] : 0xFFFFFFFFu;
    float v[kDiagVals];
#pragma unroll
    for (int q = 0; q < (int)kDiagVals; ++q) v[q] = 0.f;
    if (valid) {
        float J[3][6], res[3], lam2[3];
        const int rows = term_rows(b, c, pen, b.n2t_index[e], key, J, res, lam2);
        const float4 g = f.nodes[key];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            if (r < rows) {
                to_node_frame(J[r], g);
                int o = 0;
#pragma unroll
                for (int a = 0; a < 6; ++a)
#pragma unroll
                    for (int k2 = a; k2 < 6; ++k2) v[o++] += lam2[r] * J[r][a] * J[r][k2];
#pragma unroll
                for (int a = 0; a < 6; ++a) v[21 + a] -= lam2[r] * J[r][a] * res[r];
            }
        }
    }
    uint32_t ta = 0, tk = 0;
    if (lane < 21) triu_index(lane, ta, tk);
    chunk_runs(v, key, valid, lane, lds, [&](uint32_t rkey, uint32_t first, uint32_t last, float sum) {
        if (lane >= kDiagVals) return;
        const bool open_head = first == 0 && prev_key == rkey;
        const bool open_tail = last == 63 && next_key == rkey;
        if (open_head || open_tail) {
            b.part_diag[((size_t)ch * 2 + (first == 0 ? 0 : 1)) * kDiagVals + lane] = sum;
        } else if (lane < 21) {
            float* D = b.diag + 36 * (size_t)rkey;
            D[6 * ta + tk] = sum;
            D[6 * tk + ta] = sum;
        } else {
            b.jtr[6 * (size_t)rkey + lane - 21] = sum;
        }
    });
}

__device__ inline void pair_chunk(const Buffers& b, const Counts& c, const Penalty& pen, uint32_t ch, uint32_t lane,
                                  float* lds) {
    const FrameDev& f = *b.frame;
    const uint32_t total = c.kv2;
    const uint32_t e = ch * 64 + lane;
    const bool valid = e < total;
    const uint32_t key = valid ? b.kv2_pair[e] : 0xFFFFFFFFu;
    const uint32_t prev_key = ch > 0 ? b.kv2_pair[ch * 64 - 1] : 0xFFFFFFFFu;
    const uint32_t next_key = ch * 64 + 64 < total ? b.kv2_pair[ch * 64 + 64] : 0xFFFFFFFFu;
    float v[kPairVals];
#pragma unroll
    for (int q = 0; q < (int)kPairVals; ++q) v[q] = 0.f;
    if (valid) {
        const uint32_t pk = b.pair_key[key];
        const uint32_t i = pk >> 16, j = pk & 0xFFFF;
        const uint32_t t = b.pair_terms[e];
        float Ji[3][6], Jj[3][6], li[3];
        const int rows = pair_term_rows(b, c, pen, t, i, j, Ji, Jj, li);
        const float4 gi = f.nodes[i], gj = f.nodes[j];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            if (r < rows) {
                to_node_frame(Ji[r], gi);
                to_node_frame(Jj[r], gj);
#pragma unroll
                for (int a = 0; a < 6; ++a)
#pragma unroll
                    for (int k2 = 0; k2 < 6; ++k2) v[6 * a + k2] += li[r] * Ji[r][a] * Jj[r][k2];
            }
        }
    }
    chunk_runs(v, key, valid, lane, lds, [&](uint32_t rkey, uint32_t first, uint32_t last, float sum) {
        if (lane >= kPairVals) return;
        const bool open = (first == 0 && prev_key == rkey) || (last == 63 && next_key == rkey);
        float* out = open ? b.part_pair + ((size_t)ch * 2 + (first == 0 ? 0 : 1)) * kPairVals
                          : b.offdiag + 36 * (size_t)rkey;
        out[lane] = sum;
    });
}

__global__ void __launch_bounds__(256) assemble_chunks_kernel(const Buffers* __restrict__ bs, Penalty pen) {
    const Buffers b = frame_buffers(bs);
    const Counts& c = *b.counts;
    if (c.error) return;
    const uint32_t lane = threadIdx.x & 63;
    __shared__ float red[4 * 64 * kRedStride];
    float* lds = red + (threadIdx.x >> 6) * 64 * kRedStride;
    const uint32_t c1 = (c.kv1 + 63) / 64, c2 = (c.kv2 + 63) / 64;
    // groups of 4 consecutive chunks (one per wave), XCD-aware: an XCD works
    // through a contiguous range of node rows / pairs and keeps their term
    // records in its L2
    const uint32_t groups = (c1 + c2 + 3) / 4;
    for (uint32_t v = blockIdx.x; v < xcd_span(gridDim.x, groups); v += gridDim.x) {
        const uint32_t grp = xcd_remap(v, gridDim.x, groups);
        if (grp == 0xFFFFFFFFu) continue;
        const uint32_t w = grp * 4 + (threadIdx.x >> 6);
        if (w < c1) diag_chunk(b, c, pen, w, lane, lds);
        else if (w < c1 + c2) pair_chunk(b, c, pen, w - c1, lane, lds);
    }
}

// Runs cut by chunk boundaries: the run's first chunk holds it in slot B, or
// in slot A when the run starts at that chunk's lane 0; every later chunk in
// slot A.  Then the node rows get the damping and their block-Jacobi inverse
// (blockDiagonalInverseKernel, BlockDiagonalPreconditionerInverse.cu:6-41).
template <typename InvT>
__global__ void __launch_bounds__(256) assemble_fixup_kernel(const Buffers* __restrict__ bs) {
    const Buffers b = frame_buffers(bs);
    const FrameDev& f = *b.frame;
    const Counts& c = *b.counts;
    if (c.error) return;
    const uint32_t N = f.N, P = c.pairs;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < N + P; x += stride) {
        if (x < N) {
            const uint32_t n = x;
            const uint32_t beg = b.n2t_offset[n], end = b.n2t_offset[n + 1];
            float D[36], jtr[6];
            if (beg == end) {
#pragma unroll
                for (int q = 0; q < 36; ++q) D[q] = 0.f;
#pragma unroll
                for (int q = 0; q < 6; ++q) jtr[q] = 0.f;
            } else if (beg / 64 == (end - 1) / 64) {
#pragma unroll
                for (int q = 0; q < 36; ++q) D[q] = b.diag[36 * (size_t)n + q];
#pragma unroll
                for (int q = 0; q < 6; ++q) jtr[q] = b.jtr[6 * (size_t)n + q];
            } else {
                float v[kDiagVals];
#pragma unroll
                for (int q = 0; q < (int)kDiagVals; ++q) v[q] = 0.f;
                for (uint32_t ch = beg / 64; ch <= (end - 1) / 64; ++ch) {
                    const uint32_t slot = (ch == beg / 64 && beg % 64 != 0) ? 1u : 0u;
                    const float* p = b.part_diag + ((size_t)ch * 2 + slot) * kDiagVals;
#pragma unroll
                    for (int q = 0; q < (int)kDiagVals; ++q) v[q] += p[q];
                }
                int o = 0;
#pragma unroll
                for (int a = 0; a < 6; ++a)
#pragma unroll
                    for (int k2 = a; k2 < 6; ++k2) {
                        D[6 * a + k2] = v[o];
                        D[6 * k2 + a] = v[o];
                        ++o;
                    }
#pragma unroll
                for (int q = 0; q < 6; ++q) jtr[q] = v[21 + q];
            }
            add_node_frame_damping(D, f.nodes[n]);
#pragma unroll
            for (int q = 0; q < 36; ++q) b.diag[36 * (size_t)n + q] = D[q];
#pragma unroll
            for (int q = 0; q < 6; ++q) b.jtr[6 * (size_t)n + q] = jtr[q];
            spd_inverse6<InvT>(D, b.diag_inv + 36 * (size_t)n,
                               sizeof(InvT) == 8 ? b.diag_inv64 + 36 * (size_t)n : nullptr,
                               sizeof(InvT) == 8 ? b.linv64 + kTri * (size_t)n : nullptr);
        } else {
            const uint32_t p = x - N;
            const uint2 r = b.pair_range[p];
            if (r.x == r.y) {
#pragma unroll
                for (int q = 0; q < 36; ++q) b.offdiag[36 * (size_t)p + q] = 0.f;
                continue;
            }
            if (r.x / 64 == (r.y - 1) / 64) continue;  // written by its chunk
            float v[kPairVals];
#pragma unroll
            for (int q = 0; q < (int)kPairVals; ++q) v[q] = 0.f;
            for (uint32_t ch = r.x / 64; ch <= (r.y - 1) / 64; ++ch) {
                const uint32_t slot = (ch == r.x / 64 && r.x % 64 != 0) ? 1u : 0u;
                const float* pp = b.part_pair + ((size_t)ch * 2 + slot) * kPairVals;
#pragma unroll
                for (int q = 0; q < (int)kPairVals; ++q) v[q] += pp[q];
            }
#pragma unroll
            for (int q = 0; q < 36; ++q) b.offdiag[36 * (size_t)p + q] = v[q];
        }
    }
}

// ------------------------------------------- row Gram assembly (mixed mode)
// For the knn-weighted terms (dense, foreground, density, feature, cross) the
// Jacobian rows of node n are the term's rows J scaled by the node's knn slot
// weight w_n.  In node-centred coordinates the rows of node i are w_i u_i,
// u_i = (J_w + J_v x g_i, J_v), and those of node j are w_j u_i M_ij with
//   M_ij: (x_w, x_v) -> (x_w + x_v x (g_j - g_i), x_v)   (a change of centre).
// Every block of row i is therefore a weighted sum of the same per-term values
// G_t = (lambda^2 u_i^T u_i [21 upper-triangle entries], -lambda^2 u_i^T r [6]):
//   D_i  = sum_t w_i^2 G_t[0:21]          (PreconditionerBlocked.cu:172-284)
//   b_i  = sum_t w_i   G_t[21:27]         (JtDotResidualIndexed.cu:116-224)
//   S_ij = sum_t w_i w_j G_t[0:21],  B_ij = S_ij M_ij
//                                         (JtJMaterializer.cu:270-365)
// where w_j sums over the knn slots holding j (the reference's pair
// multiplicity) and each Node2Term entry of i contributes once (its slot
// multiplicity).  For row i that is the product G^T W of the row's terms (K)
// with a weight matrix W whose columns are the diagonal (w_i^2), J^T r (w_i)
// and the upper partners j > i (w_i w_j) -- run on the matrix cores with
// v_mfma_f32_16x16x4_f32 (FP32 products and accumulation), four terms per
// instruction, two 16-row tiles for the 27 values, one 16-column tile per 16
// weight columns.  Graph-edge (smooth) terms have different rows per node: they
// enter the diagonal and J^T r here and their pair blocks in the fix-up.
// One wave per row and 64 Node2Term entries per step of the row: the lanes
// evaluate one entry each and stage G_t and the entry's weight row in LDS; the
// wave then runs the product.  Each pass covers 16 weight columns (the
// diagonal, J^T r and 14 partners); rows with more upper partners take further
// passes.  The partner columns of every knn slot come from the index
// (n2t_col); no NodePair2Term lists are read.
constexpr uint32_t kGramVals = 27;
constexpr uint32_t kGramCols = 16;
constexpr uint16_t kNoNode = 0xFFFF;

typedef float gram_f4 __attribute__((ext_vector_type(4)));

// A term record as loaded: 64 B (dense, foreground, feature, cross) or 32 B
// (graph edge; the upper half is not loaded and not read).  Loading it apart
// from its use lets a row fetch its next 64 entries while the current ones are
// on the matrix cores.  t is always a valid term: lanes past a row's end load
// the row's last entry again (and stage it with zero weights).
struct RawRec {
    uint4 q[4];
};

__device__ inline RawRec load_record(const Buffers& b, const TermBounds& c, uint32_t t) {
    RawRec r;
    uint32_t k;
    const int type = type_of(c, t, k);
    const uint4* p = type == kSmooth ? reinterpret_cast<const uint4*>(b.smooth + k)
                   : type == kDense  ? reinterpret_cast<const uint4*>(b.dense + k)
                   : type == kForeground ? reinterpret_cast<const uint4*>(b.fg + k)
                   : type == kFeature ? reinterpret_cast<const uint4*>(b.feat + k)
                                      : reinterpret_cast<const uint4*>(b.cross + k);
    r.q[0] = p[0];
    r.q[1] = p[1];
    if (type != kSmooth) {
        r.q[2] = p[2];
        r.q[3] = p[3];
    }
    return r;
}

// lambda^2 of every term type (registers, not a by-value struct read through
// a selected address).
struct Lambda2 {
    float smooth, density, foreground, feature, cross;
    __device__ explicit Lambda2(const Penalty& p)
        : smooth(p.smooth * p.smooth), density(p.density * p.density), foreground(p.foreground * p.foreground),
          feature(p.feature * p.feature), cross(p.cross * p.cross) {}
};

// LDS rows of one staged entry, 16 words each: a G_t row (values 0..15 for
// the first product, then values 16..26 and zeros for the second) and the
// weight row.  Word c of row r sits at 16 r + (c ^ swz(r)),
// swz(r) = 2 ((r >> 1) & 7): a lane's float2 stores of its own row and the
// product's column reads (4 rows x 16 columns per step) both fall on distinct
// banks.  The two halves of G_t share one buffer (8 KB of LDS per row wave).
constexpr uint32_t kRowWords = 16;
__device__ inline uint32_t swz(uint32_t r) { return 2 * ((r >> 1) & 7); }

// G_t of `rows` Jacobian rows (world frame; moved to the node frame of g here)
// with their lambda^2 and residuals.  Row 0 always exists; rows 1, 2 are added
// when present.
template <int kMaxRows>
__device__ inline void rows_gram(float J[3][6], const float res[3], const float lam2[3], int rows, const float4 g,
                                 float (&G)[kGramVals]) {
    to_node_frame(J[0], g);
    {
        int o = 0;
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            const float la = lam2[0] * J[0][a];
#pragma unroll
            for (int k2 = a; k2 < 6; ++k2) G[o++] = la * J[0][k2];
            G[21 + a] = -la * res[0];
        }
    }
#pragma unroll
    for (int r = 1; r < kMaxRows; ++r) {
        if (r < rows) {
            to_node_frame(J[r], g);
            int o = 0;
#pragma unroll
            for (int a = 0; a < 6; ++a) {
                const float la = lam2[r] * J[r][a];
#pragma unroll
                for (int k2 = a; k2 < 6; ++k2) G[o++] += la * J[r][k2];
                G[21 + a] -= la * res[r];
            }
        }
    }
}

// G_t of a point term (graph edge, feature, cross view) in closed form.  Its
// rows are h_r (q x e_r, e_r) in node-centred coordinates, q = p - g_i the
// term's point relative to the node, with residuals rho_r.  With
// H_r = lambda^2 h_r^2 and V_r = lambda^2 h_r rho_r:
//   rotation block    [q]x H [q]x^T   (sum_r H_r (q x e_r)(q x e_r)^T)
//   rotation / trans  [q]x H
//   translation       diag(H)
//   -J^T r            -(q x V, V)
// q is formed as one difference of nearby points, so the lever arm keeps its
// FP32 precision.
__device__ inline void point_gram(float qx, float qy, float qz, float H0, float H1, float H2, float V0, float V1,
                                  float V2, float (&G)[kGramVals]) {
    const float xx = qx * qx, yy = qy * qy, zz = qz * qz;
    G[0] = zz * H1 + yy * H2;  // M00
    G[1] = -qx * qy * H2;      // M01
    G[2] = -qx * qz * H1;      // M02
    G[3] = 0.f;                // C00
    G[4] = -qz * H1;           // C01
    G[5] = qy * H2;            // C02
    G[6] = zz * H0 + xx * H2;  // M11
    G[7] = -qy * qz * H0;      // M12
    G[8] = qz * H0;            // C10
    G[9] = 0.f;                // C11
    G[10] = -qx * H2;          // C12
    G[11] = yy * H0 + xx * H1; // M22
    G[12] = -qy * H0;          // C20
    G[13] = qx * H1;           // C21
    G[14] = 0.f;               // C22
    G[15] = H0;
    G[16] = 0.f;
    G[17] = 0.f;
    G[18] = H1;
    G[19] = 0.f;
    G[20] = H2;
    G[21] = -(qy * V2 - qz * V1);
    G[22] = -(qz * V0 - qx * V2);
    G[23] = -(qx * V1 - qy * V0);
    G[24] = -V0;
    G[25] = -V1;
    G[26] = -V2;
}

// One half of G_t into the lane's LDS row: values 0..15, or 16..26 and zeros.
template <int kHalf>
__device__ inline void store_half(const float (&G)[kGramVals], float* gl) {
    const uint32_t f = swz(threadIdx.x);
    if (kHalf == 0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) *reinterpret_cast<float2*>(gl + ((2 * q) ^ f)) = make_float2(G[2 * q], G[2 * q + 1]);
    } else {
#pragma unroll
        for (int q = 0; q < 5; ++q)
            *reinterpret_cast<float2*>(gl + ((2 * q) ^ f)) = make_float2(G[16 + 2 * q], G[17 + 2 * q]);
        *reinterpret_cast<float2*>(gl + (10 ^ f)) = make_float2(G[26], 0.f);
#pragma unroll
        for (int q = 6; q < 8; ++q) *reinterpret_cast<float2*>(gl + ((2 * q) ^ f)) = make_float2(0.f, 0.f);
    }
}

// Partner columns of row i held by the index or searched in the row's pair keys.
struct GramColumns {
    uint32_t c0;       // first weight column of this pass
    uint32_t ubeg, U;  // the row's upper pairs
    int mode;          // 0: no partner columns, 1: n2t_col, 2: pair-key search
};

// The weight row of a knn-weighted term for row node i: diagonal w_i^2, J^T r
// w_i, and per upper partner j the weight w_i w_j summed over the knn slots
// holding j (every such slot stores the same sum).  Slots outside this pass
// store into the padding column.
__device__ inline void store_knn_weights(const Buffers& b, const GramColumns& gc, ushort4 knn, float4 w, uint32_t i,
                                         uint32_t pc, float* wl, float* wdummy) {
    const float wi = slot_weight(knn, w, i);
    const uint32_t f = swz(threadIdx.x);
    if (gc.c0 == 0) {
        wl[0 ^ f] = wi * wi;
        wl[1 ^ f] = wi;
    }
    if (gc.mode == 0) return;
    const uint32_t n[4] = {knn.x, knn.y, knn.z, knn.w};
    const bool distinct = knn.x != knn.y && knn.x != knn.z && knn.x != knn.w && knn.y != knn.z && knn.y != knn.w &&
                          knn.z != knn.w;
    float wq[4] = {w.x, w.y, w.z, w.w};  // distinct knn: the slot rule gives knn[q] the weight w[q]
    if (!distinct) {
#pragma unroll
        for (int q = 0; q < 4; ++q) wq[q] = slot_weight(knn, w, n[q]);
    }
    uint32_t col[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const bool up = n[q] > i && n[q] != kNoNode;
        wq[q] = up ? wi * wq[q] : 0.f;
        col[q] = (pc >> (8 * q)) & 0xFFu;
    }
    if (gc.mode == 2) {
        // more partners than the index's u8 columns: search the row's pair keys
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            col[q] = 0xFFu;
            if (wq[q] == 0.f) continue;
            const uint32_t key = (i << 16) | n[q];
            uint32_t lo = 0, hi = gc.U;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (b.pair_key[gc.ubeg + mid] < key) lo = mid + 1;
                else hi = mid;
            }
            col[q] = 2 + lo;
        }
    }
    float sum[4] = {wq[0], wq[1], wq[2], wq[3]};
    if (!distinct) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            sum[q] = 0.f;
#pragma unroll
            for (int qq = 0; qq < 4; ++qq) sum[q] += col[qq] == col[q] ? wq[qq] : 0.f;
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t r = col[q] - gc.c0;
        *(wq[q] != 0.f && r < kGramCols ? wl + (r ^ f) : wdummy) = sum[q];
    }
}

// One Node2Term entry of row i (term t, record raw): its weight row stored to
// LDS, its G_t returned in registers (zeros for an empty lane).
__device__ inline void stage_entry(const Buffers& b, const TermBounds& c, const Lambda2& l2, bool density,
                                   const GramColumns& gc, uint32_t t, bool live, const RawRec& raw, uint32_t pc,
                                   uint32_t i, const float4 g, float (&G)[kGramVals], float* wl, float* wdummy) {
    // zero row (swizzled positions: the lanes' stores stay on distinct banks)
    const uint32_t fz = swz(threadIdx.x);
#pragma unroll
    for (int q = 0; q < 8; ++q) *reinterpret_cast<float2*>(wl + ((2 * q) ^ fz)) = make_float2(0.f, 0.f);
    uint32_t k;
    const int type = type_of(c, t, k);
    if (type == kSmooth) {
        // graph edge (computeSmoothTermJacobian, geometry_icp_jacobian.cuh:323-343):
        // rows of node i or j, unit weight, no partner columns here
        SmoothRec sm;
        __builtin_memcpy(&sm, &raw, sizeof(SmoothRec));
        const bool node_i = i == (sm.ij & 0xFFFF);
        const float on = sm.valid ? l2.smooth : 0.f;
        const float sv = node_i ? on : -on;
        const float qx = (node_i ? sm.Ti_xj[0] : sm.Ti_xj[0] - sm.r[0]) - g.x;
        const float qy = (node_i ? sm.Ti_xj[1] : sm.Ti_xj[1] - sm.r[1]) - g.y;
        const float qz = (node_i ? sm.Ti_xj[2] : sm.Ti_xj[2] - sm.r[2]) - g.z;
        point_gram(qx, qy, qz, on, on, on, sv * sm.r[0], sv * sm.r[1], sv * sm.r[2], G);
        if (gc.c0 == 0) {
            const uint32_t f = swz(threadIdx.x);
            const float one = live ? 1.f : 0.f;
            wl[0 ^ f] = one;
            wl[1 ^ f] = one;
        }
        return;
    }
    ushort4 knn;
    float4 w;
    if (type == kDense || type == kForeground) {
        ScalarRec rec;
        __builtin_memcpy(&rec, &raw, sizeof(ScalarRec));
        knn = rec.knn;
        w = live ? rec.w : make_float4(0.f, 0.f, 0.f, 0.f);
        float J[3][6], res[3], lam2[3];
#pragma unroll
        for (int q = 0; q < 6; ++q) J[0][q] = rec.J[q];
        res[0] = rec.r;
        lam2[0] = type == kDense ? 1.0f : l2.foreground;
        int rows = 1;
        if (type == kDense && density) {
            const ScalarRec& d = b.density[k];
#pragma unroll
            for (int q = 0; q < 6; ++q) J[1][q] = d.J[q];
            res[1] = d.r;
            lam2[1] = l2.density;
            rows = 2;
        }
        rows_gram<2>(J, res, lam2, rows, g, G);
    } else {
        // point-to-point rows (computePointToPointICPTermJacobian,
        // geometry_icp_jacobian.cuh:91-117): h_r = huber(r_r), rho_r = h_r r_r
        P2PRec rec;
        __builtin_memcpy(&rec, &raw, sizeof(P2PRec));
        knn = rec.knn;
        w = live ? rec.w : make_float4(0.f, 0.f, 0.f, 0.f);
        // bit select: a select of two fields would become a load through a
        // selected address (and the struct a scratch array)
        const uint32_t mask = type == kFeature ? 0xFFFFFFFFu : 0u;
        const float l = __uint_as_float((__float_as_uint(l2.feature) & mask) | (__float_as_uint(l2.cross) & ~mask));
        const float4 p = rec.warped;
        const float r0 = p.x - rec.target.x, r1 = p.y - rec.target.y, r2 = p.z - rec.target.z;
        const float h0 = huber(r0), h1 = huber(r1), h2 = huber(r2);
        const float H0 = l * (h0 * h0), H1 = l * (h1 * h1), H2 = l * (h2 * h2);
        point_gram(p.x - g.x, p.y - g.y, p.z - g.z, H0, H1, H2, H0 * r0, H1 * r1, H2 * r2, G);
    }
    store_knn_weights(b, gc, knn, w, i, pc, wl, wdummy);
}

// A step whose entries are all dense terms (the bulk of every row), partner
// columns from the index, no density rows: no type dispatch.  Same values as
// stage_entry.
__device__ inline void stage_dense(const Buffers& b, const TermBounds& c, const Lambda2& l2, const GramColumns& gc,
                                   uint32_t t, bool live, const RawRec& raw, uint32_t pc, uint32_t i, const float4 g,
                                   float (&G)[kGramVals], float* wl, float* wdummy) {
    const uint32_t f = swz(threadIdx.x);
#pragma unroll
    for (int q = 0; q < 8; ++q) *reinterpret_cast<float2*>(wl + ((2 * q) ^ f)) = make_float2(0.f, 0.f);
    ScalarRec rec;
    __builtin_memcpy(&rec, &raw, sizeof(ScalarRec));
    float J[3][6], res[3], lam2[3] = {1.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < 6; ++q) J[0][q] = rec.J[q];
    res[0] = rec.r;
    rows_gram<1>(J, res, lam2, 1, g, G);
    const ushort4 knn = rec.knn;
    const float4 w = live ? rec.w : make_float4(0.f, 0.f, 0.f, 0.f);
    const bool distinct = knn.x != knn.y && knn.x != knn.z && knn.x != knn.w && knn.y != knn.z && knn.y != knn.w &&
                          knn.z != knn.w;
    if (!distinct) {  // repeated knn nodes: the general rule
        store_knn_weights(b, gc, knn, w, i, pc, wl, wdummy);
        return;
    }
    const float wi = slot_weight(knn, w, i);
    if (gc.c0 == 0) {
        wl[0 ^ f] = wi * wi;
        wl[1 ^ f] = wi;
    }
    // distinct knn: slot q holds knn[q] with weight w[q]; its column is 0xFF
    // unless knn[q] is an upper partner of i
    const float wq[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t r = ((pc >> (8 * q)) & 0xFFu) - gc.c0;
        if (r < kGramCols) wl[r ^ f] = wi * wq[q];
    }
}

// acc += G^T W over the staged entries: `steps` products of 4 entries, the
// operands of lane (m, kq) at 64 u + roff[u & 3].
__device__ inline void gram_product(const float* Gs, const float* Ws, const uint32_t (&roff)[4], uint32_t steps,
                                    gram_f4& acc) {
    uint32_t st = 0;
    if (steps == 16) {  // a full step: every operand read is issued before the products
        float a[16], bv[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const uint32_t o = 4 * kRowWords * u + roff[u & 3];
            a[u] = Gs[o];
            bv[u] = Ws[o];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], bv[u], acc, 0, 0, 0);
        return;
    }
    for (; st + 4 <= steps; st += 4) {  // four steps' operands in flight at once
        float a[4], bv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t o = 4 * kRowWords * (st + u) + roff[u];  // st is a multiple of 4
            a[u] = Gs[o];
            bv[u] = Ws[o];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], bv[u], acc, 0, 0, 0);
    }
    for (; st < steps; ++st) {
        const uint32_t j = st & 3;
        const uint32_t o = 4 * kRowWords * st + (j == 0 ? roff[0] : j == 1 ? roff[1] : j == 2 ? roff[2] : roff[3]);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(Gs[o], Ws[o], acc, 0, 0, 0);
    }
}

// The columns of one finished pass of row i: lane (m, kq) holds values
// 4 kq + r (lo) and 16 + 4 kq + r (hi) of column c0 + m.
__device__ inline void store_columns(const Buffers& b, uint32_t i, uint32_t c0, uint32_t cols, uint32_t ubeg,
                                     const gram_f4& lo, const gram_f4& hi, uint32_t m, uint32_t kq) {
    const uint32_t n = c0 + m;
    if (n == 0) {
        float* D = b.diag + 36 * (size_t)i;  // 32 values; the fix-up reads the first 21
        *reinterpret_cast<float4*>(D + 4 * kq) = make_float4(lo[0], lo[1], lo[2], lo[3]);
        *reinterpret_cast<float4*>(D + 16 + 4 * kq) = make_float4(hi[0], hi[1], hi[2], hi[3]);
    } else if (n == 1) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t q = 16 + 4 * kq + r;
            if (q >= 21 && q < kGramVals) b.jtr[6 * (size_t)i + q - 21] = hi[r];
        }
    } else if (n < cols) {
        float* S = b.offdiag + 36 * (size_t)(ubeg + n - 2);  // 24 values; the fix-up reads 21
        *reinterpret_cast<float4*>(S + 4 * kq) = make_float4(lo[0], lo[1], lo[2], lo[3]);
        if (kq < 2) *reinterpret_cast<float4*>(S + 16 + 4 * kq) = make_float4(hi[0], hi[1], hi[2], hi[3]);
    }
}

__device__ inline uint32_t lane_u32(uint32_t v, uint32_t k) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)k);
}
__device__ inline float lane_f32(float v, uint32_t k) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)k));
}

// Rows per row wave (kRows: 1, 2 or 4, a launch shape).  A wave runs the
// Node2Term entries of kRows consecutive rows (and of each row's column
// passes) as one stream of 64-entry steps: the entries and term records of a
// row's first step are fetched while the previous row's last step computes,
// so a row start costs no dependent loads of its own.  Every row's products
// run in the same order for any kRows.

// Position in a wave's stream: row k (of the wave), column pass c0, step e0.
struct GramCursor {
    uint32_t k, c0, e0;
};

// diag_only: the diagonal blocks and J^T r alone (the matrix-free solve).
template <uint32_t kGramRows>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4)))
gram_rows_kernel(const Buffers* __restrict__ bs, Penalty pen, int diag_only) {
    const Buffers b = frame_buffers(bs);
    const uint32_t lane = threadIdx.x;
    // XCD-aware rows: each XCD takes a contiguous range of the frame's rows
    // (whose entries share term records) into its L2, spread over all XCDs
    // whatever the frame's node count; lane k loads row r0 + k's offsets
    // (clamped to the capacity; rows past the frame's node count are not used).
    const FrameDev& f = *b.frame;
    const Counts& c = *b.counts;
    const uint32_t r0 = xcd_remap(blockIdx.x, gridDim.x, (f.N + kGramRows - 1) / kGramRows) * kGramRows;
    if (c.error || r0 >= f.N) return;
    const uint32_t ol = min(r0 + lane, b.cap_nodes);
    const uint32_t off_l = b.n2t_offset[ol];
    const uint32_t poff_l = diag_only ? 0u : b.pair_offset[ol];
    const uint32_t nrows = min(kGramRows, f.N - r0);
    const float4 g_l = f.nodes[min(r0 + lane, f.N - 1)];
    __shared__ float Gs[64 * kRowWords];
    __shared__ float Wl[64 * kRowWords];
    __shared__ float Wd[64];  // weight stores of slots outside the pass
    const uint32_t m = lane & 15, kq = lane >> 4;
    float* gl = Gs + kRowWords * lane;
    float* wl = Wl + kRowWords * lane;
    float* wd = Wd + lane;
    // read offsets of lane (m, kq): row 4 u + kq, word m ^ swz(row); swz
    // depends on u only through u mod 4
    uint32_t roff[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) roff[j] = kRowWords * kq + (m ^ swz(4 * j + kq));
    const Lambda2 l2(pen);
    const bool density = b.density && pen.density != 0.f;
    const TermBounds tb(c);
    const gram_f4 zero4 = {0.f, 0.f, 0.f, 0.f};

    // rows without entries: zero diagonal and J^T r (no pairs)
    for (uint32_t k = 0; k < nrows; ++k)
        if (lane_u32(off_l, k) == lane_u32(off_l, k + 1)) store_columns(b, r0 + k, 0, 2, 0, zero4, zero4, m, kq);

    // the stream: every (row, pass) of the wave's rows that has entries
    auto row_beg = [&](uint32_t k) { return lane_u32(off_l, k); };
    auto row_end = [&](uint32_t k) { return lane_u32(off_l, k + 1); };
    auto row_cols = [&](uint32_t k) { return 2 + lane_u32(poff_l, k + 1) - lane_u32(poff_l, k); };
    auto skip_empty = [&](GramCursor& q) {
        while (q.k < nrows && row_beg(q.k) == row_end(q.k)) ++q.k;
        q.c0 = 0;
        q.e0 = q.k < nrows ? row_beg(q.k) : 0;
    };
    auto advance = [&](GramCursor q) {
        q.e0 += 64;
        if (q.e0 < row_end(q.k)) return q;
        q.c0 += kGramCols;
        if (q.c0 < row_cols(q.k)) {
            q.e0 = row_beg(q.k);
            return q;
        }
        ++q.k;
        skip_empty(q);
        return q;
    };
    GramCursor cur{0, 0, 0};
    skip_empty(cur);
    if (cur.k >= nrows) return;
    // two-stage fetch: the records of the next step and the entries of the step
    // after it are in flight while a step computes.  Entry indices past a row's
    // end are clamped to its last entry (staged with zero weights).
    uint32_t e = min(cur.e0 + lane, row_end(cur.k) - 1);
    // n2t_col exists only with partner columns (it is not allocated for the
    // matrix-free solve, which assembles the diagonal alone)
    const bool cols_in = !diag_only;
    uint32_t t_cur = b.n2t_index[e];
    uint32_t pc_cur = cols_in ? b.n2t_col[e] : 0xFFFFFFFFu;
    RawRec rec_cur = load_record(b, tb, t_cur);
    GramCursor nxt = advance(cur);
    uint32_t t_nxt = t_cur, pc_nxt = pc_cur;
    if (nxt.k < nrows) {
        e = min(nxt.e0 + lane, row_end(nxt.k) - 1);
        t_nxt = b.n2t_index[e];
        pc_nxt = cols_in ? b.n2t_col[e] : 0xFFFFFFFFu;
    }
    gram_f4 lo = zero4, hi = zero4;
    while (cur.k < nrows) {
        const uint32_t k = cur.k, i = r0 + k;
        const uint32_t end = row_end(k), cols = diag_only ? 2u : row_cols(k);
        const uint32_t ubeg = lane_u32(poff_l, k);
        // partner columns held by n2t_col while they fit its u8 columns
        const int mode = diag_only ? 0 : (cols < 0xFF ? 1 : 2);
        const GramColumns gc{cur.c0, ubeg, cols - 2, mode};
        const float4 g = make_float4(lane_f32(g_l.x, k), lane_f32(g_l.y, k), lane_f32(g_l.z, k), 0.f);
        const uint32_t pc = mode == 1 ? pc_cur : 0xFFFFFFFFu;
        const bool live = cur.e0 + lane < end;
        float G[kGramVals];
        // entries ascend by term within a row and lane 63 holds the step's
        // last one: the step is all dense when that one is
        const bool dense_step = mode == 1 && !density && lane_u32(t_cur, 63) < tb.o0;
        if (dense_step)
            stage_dense(b, tb, l2, gc, t_cur, live, rec_cur, pc, i, g, G, wl, wd);
        else
            stage_entry(b, tb, l2, density, gc, t_cur, live, rec_cur, pc, i, g, G, wl, wd);
        store_half<0>(G, gl);
        const GramCursor nn = advance(nxt);
        RawRec rec_nxt = rec_cur;
        if (nxt.k < nrows) {
            if (lane_u32(t_nxt, 63) < tb.o0) {
                const uint4* rp = reinterpret_cast<const uint4*>(b.dense + t_nxt);
#pragma unroll
                for (int q = 0; q < 4; ++q) rec_nxt.q[q] = rp[q];
            } else {
                rec_nxt = load_record(b, tb, t_nxt);
            }
        }
        uint32_t t_nn = t_nxt, pc_nn = pc_nxt;
        if (nn.k < nrows) {
            const uint32_t enn = min(nn.e0 + lane, row_end(nn.k) - 1);
            t_nn = b.n2t_index[enn];
            pc_nn = cols_in ? b.n2t_col[enn] : 0xFFFFFFFFu;
        }
        const uint32_t steps = (min(end - cur.e0, 64u) + 3) / 4;
        wave_lds_sync();
        gram_product(Gs, Wl, roff, steps, lo);
        wave_lds_sync();  // the G rows are rewritten with the second half
        store_half<1>(G, gl);
        wave_lds_sync();
        gram_product(Gs, Wl, roff, steps, hi);
        wave_lds_sync();  // LDS is rewritten by the next step
        if (nxt.k != k || nxt.c0 != cur.c0) {  // the pass ends with this step
            store_columns(b, i, cur.c0, cols, ubeg, lo, hi, m, kq);
            lo = zero4;
            hi = zero4;
        }
        cur = nxt;
        nxt = nn;
        t_cur = t_nxt;
        pc_cur = pc_nxt;
        rec_cur = rec_nxt;
        t_nxt = t_nn;
        pc_nxt = pc_nn;
    }
}

// Fix-up of the row Gram assembly: node rows get the damping and their
// block-Jacobi inverse with the FP64 Cholesky factor
// (blockDiagonalInverseKernel, BlockDiagonalPreconditionerInverse.cu:6-41).
// The pair blocks are finished by the transform that starts the CG
// (cg_transform_kernel), which needs both nodes' factors.
// inverse64: also the FP64 inverse (the matrix-free solve's preconditioner);
// the fused CG needs only the factors.  full: also the damped blocks (diag,
// in place of the Gram sums) and their FP32 inverses (diag_inv); without it
// the Gram sums stay in diag, and a later call with full forms both (inspection
// and the materialized SpMV stage).
__global__ void __launch_bounds__(256) gram_fixup_kernel(const Buffers* __restrict__ bs, int inverse64, int full) {
    const Buffers b = frame_buffers(bs);
    const FrameDev& f = *b.frame;
    const Counts& c = *b.counts;
    if (c.error) return;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < f.N; n += stride) {
        const float* Dn = b.diag + 36 * (size_t)n;
        float v[21];
#pragma unroll
        for (int q = 0; q < 21; ++q) v[q] = Dn[q];
        float D[36];
        int o = 0;
#pragma unroll
        for (int a = 0; a < 6; ++a)
#pragma unroll
            for (int k2 = a; k2 < 6; ++k2) {
                D[6 * a + k2] = v[o];
                D[6 * k2 + a] = v[o];
                ++o;
            }
        add_node_frame_damping(D, f.nodes[n]);
        if (full) {
#pragma unroll
            for (int q = 0; q < 36; ++q) b.diag[36 * (size_t)n + q] = D[q];
        }
        spd_inverse6<double>(D, full ? b.diag_inv + 36 * (size_t)n : nullptr,
                             inverse64 ? b.diag_inv64 + 36 * (size_t)n : nullptr, b.linv64 + kTri * (size_t)n);
    }
}

// ---------------------------------------------------------- total residual
// computeTermResidualKernel (ResidualEvaluator.cu:71-142): sum of squared
// lambda-weighted residuals, accumulated in double.
__global__ void __launch_bounds__(256) cost_kernel(const Buffers* __restrict__ bs, Penalty pen, int density_in_jtj) {
    const Buffers b = frame_buffers(bs);
    const Counts& c = *b.counts;
    double acc = 0.0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < c.off[4] && !c.error; t += stride) {
        uint32_t k;
        const int type = type_of(c, t, k);
        if (type == kDense) {
            const double r = b.dense[k].r;
            acc += r * r;
            if (density_in_jtj && b.density) {
                const double d = (double)pen.density * b.density[k].r;
                acc += d * d;
            }
        } else if (type == kForeground) {
            const double r = (double)pen.foreground * b.fg[k].r;
            acc += r * r;
        } else if (type == kSmooth) {
            float J[3][6], res[3];
            smooth_rows(b.smooth[k], true, J, res);
            for (int ch = 0; ch < 3; ++ch) acc += (double)pen.smooth * pen.smooth * res[ch] * res[ch];
        } else {
            float J[3][6], res[3];
            p2p_rows(type == kFeature ? b.feat[k] : b.cross[k], J, res);
            const double l = type == kFeature ? pen.feature : pen.cross;
            for (int ch = 0; ch < 3; ++ch) acc += l * l * res[ch] * res[ch];
        }
    }
    acc = wave_sum(acc);
    __shared__ double part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    // one partial per block; the host sums them in block order
    if (threadIdx.x == 0) b.partial[(size_t)kSlotCost * b.partial_stride + blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// ------------------------------------------------------------- host launch
// Terms at the warp field se3[which].
// Workgroups per frame: a grid-stride loop over the terms (a tuned count from
// the launch table, else 1024).
// Rows per Gram row wave: the table's choice, else one (as fast as two in a
// 32-frame launch, with fewer frames in flight per XCD: 0.68 against 0.96 GB
// fetched per launch).
uint32_t gram_rows_for(uint32_t nodes, uint32_t frames) {
    const uint32_t tuned = launch_params(nodes, frames).gram_rows;
    return tuned == 2 || tuned == 4 ? tuned : 1u;
}

template <uint32_t kRows>
static void launch_gram_rows(const Batch& B, const Penalty& pen, bool diag_only, hipStream_t stream) {
    const uint32_t waves = ((B.nodes + kRows - 1) / kRows + kXcds - 1) / kXcds * kXcds;
    hipLaunchKernelGGL(gram_rows_kernel<kRows>, dim3(waves, B.frames), dim3(64), 0, stream, B.dev, pen,
                       diag_only ? 1 : 0);
}

// The row Gram assembly alone (assemble and the autotuner).
void launch_gram_rows_once(const Batch& B, const Penalty& pen, bool diag_only, hipStream_t stream) {
    switch (gram_rows_for(B.nodes, B.frames)) {
        case 1: launch_gram_rows<1>(B, pen, diag_only, stream); break;
        case 4: launch_gram_rows<4>(B, pen, diag_only, stream); break;
        default: launch_gram_rows<2>(B, pen, diag_only, stream); break;
    }
}

uint32_t eval_workgroups_for(uint32_t nodes, uint32_t frames) {
    const uint32_t tuned = launch_params(nodes, frames).eval_workgroups;
    return tuned ? tuned : 1024u;
}

void launch_eval_terms(const Batch& B, uint32_t which, bool fp64_geometry, hipStream_t stream) {
    const uint32_t grid = eval_workgroups_for(B.nodes, B.frames);
    if (fp64_geometry)
        hipLaunchKernelGGL(eval_terms_kernel<double>, dim3(grid, B.frames), dim3(256), 0, stream, B.dev, which);
    else
        hipLaunchKernelGGL(eval_terms_kernel<float>, dim3(grid, B.frames), dim3(256), 0, stream, B.dev, which);
}

// gram: the row Gram assembly with FP64 factors (mixed precision mode);
// otherwise the reference's structure -- segmented sums over Node2Term and
// NodePair2Term -- with FP32 inverses.
// diag_only (with gram): diagonal blocks and J^T r only, for the matrix-free solve.
void launch_assemble(const Batch& B, const Penalty& pen, bool gram, bool diag_only, hipStream_t stream) {
    const uint32_t fix_blocks = std::min<uint32_t>((8 * B.nodes + 255) / 256, 1024u);  // grid-stride over N + P
    if (gram) {
        launch_gram_rows_once(B, pen, diag_only, stream);
        hipLaunchKernelGGL(gram_fixup_kernel, dim3((B.nodes + 63) / 64, B.frames), dim3(64), 0, stream, B.dev,
                           diag_only ? 1 : 0, diag_only ? 1 : 0);
    } else {
        hipLaunchKernelGGL(assemble_chunks_kernel, dim3(2048, B.frames), dim3(256), 0, stream, B.dev, pen);
        hipLaunchKernelGGL(assemble_fixup_kernel<float>, dim3(fix_blocks, B.frames), dim3(256), 0, stream, B.dev);
    }
}

// The damped diagonal blocks and their FP32 inverses after a fused-CG
// assembly (which leaves the Gram sums in diag): once per assembly.
void launch_diag_blocks(const Batch& B, hipStream_t stream) {
    hipLaunchKernelGGL(gram_fixup_kernel, dim3((B.nodes + 63) / 64, B.frames), dim3(64), 0, stream, B.dev, 0, 1);
}

// Returns the number of partials written to slot kSlotCost (per frame).
uint32_t launch_cost(const Batch& B, const Penalty& pen, bool density_in_jtj, hipStream_t stream) {
    const uint32_t blocks = 256;
    hipLaunchKernelGGL(cost_kernel, dim3(blocks, B.frames), dim3(256), 0, stream, B.dev, pen, density_in_jtj ? 1 : 0);
    return blocks;
}

}  // namespace mvdr
