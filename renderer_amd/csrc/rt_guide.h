/*
 * rt_guide.h — learned path-guiding network: device state and launchers.
 *
 * Network: 80 input features (12-frequency encoding of the position in the
 * scene box, shading normal, outgoing direction) -> 64 -> 64 -> 64 (ReLU)
 * -> 48 outputs = 8 von Mises-Fisher lobes (mean direction, log
 * concentration, mixture logit) + the logit of the probability of sampling
 * the guide instead of the BSDF.  FP16 weights and activations, FP32
 * accumulation, FP32 master weights and Adam.
 *
 * Flat parameter layout (FP32 master, 16,624 values):
 *   W1[64][80] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[48][64] b4[48]
 *
 * FP16 copies of the weights, pre-tiled as A operands of
 * v_mfma_f32_16x16x32_f16 (16 rows x 32 K, 64 lanes x 8 halves per tile), so
 * a wave reads each operand as 64 contiguous 16-byte words:
 *
 * Forward tiles (17,408 values; inference and training), tiles (ot, kb):
 * lane m + 16 q, slot j holds W[row(ot, m)][32 kb + col(q, j)].
 *   W1 (inputs padded to 96): col = 8 q + j, the feature order itself.
 *   W2..W4: col = 4 q + j for j < 4, 16 + 4 q + (j - 4) for j >= 4, the
 *   order in which a lane holds the previous layer's outputs (its D tiles
 *   2 kb and 2 kb + 1), so a layer's result is the next one's B operand.
 *   W1..W3: row = 16 ot + m.  W4: row m of tile ot is the output that lane
 *   group m / 4 keeps in slot 4 ot + m % 4: slots 0..9 are lobes 2 g and
 *   2 g + 1 of lane group g (outputs 10 g .. 10 g + 9), slot 10 of group 0
 *   is the guide probability logit (output 40); the other slots are zero.
 *
 * Backward tiles (12,288 values; training), the transposed weights
 * W2^T, W3^T, W4^T, tiles (it, kb): lane m + 16 q, slot j holds
 * W[out(q, j, kb)][16 it + m].  W2, W3: out = 32 kb + col(q, j) as above, so
 * a layer's output gradients, held like its outputs, are the B operand.  W4:
 * out = the output of lane group q, slot 8 kb + j (zero for kb = 1, j >= 4
 * and for slots without an output), so the head's gradients, held by the
 * lanes that evaluated the head, are the B operand.
 */
#ifndef RT_GUIDE_H
#define RT_GUIDE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GD_IN 80
#define GD_H 64
#define GD_OUT 48
#define GD_LOBES 8
#define GD_NPARAM 16624
#define GD_OFF_W1 0
#define GD_OFF_B1 5120
#define GD_OFF_W2 5184
#define GD_OFF_B2 9280
#define GD_OFF_W3 9344
#define GD_OFF_B3 13440
#define GD_OFF_W4 13504
#define GD_OFF_B4 16576
#define GD_NW16I 17408   /* forward tiles */
#define GD_W16I_W1 0
#define GD_W16I_W2 6144
#define GD_W16I_W3 10240
#define GD_W16I_W4 14336
#define GD_NW16B 12288   /* backward tiles */
#define GD_W16B_W2 0
#define GD_W16B_W3 4096
#define GD_W16B_W4 8192
#define GD_NW16 (GD_NW16I + GD_NW16B)
#define GD_KLO -3.0f
#define GD_KHI 10.0f
#define GD_REC 16        /* floats per training record */
#define GD_PARAMS 48     /* halves of head output per path vertex */
#define GD_WCLAMP 100.0f /* cap of a normalised sample weight */
/* probability of sampling the guide: GD_ALPHA_LO + GD_ALPHA_SPAN sigmoid(output 40) */
#define GD_ALPHA_LO 0.05f
#define GD_ALPHA_SPAN 0.9f
#define GD_TRAIN_CHUNKS 4 /* training steps per render pass (record capacity = 4 batches) */
/* Persistent grids of the network kernels: the blocks one CU holds at once
 * (LDS: 45 KB per inference block, 155 KB per training block).  More blocks
 * would only reload the weights into LDS and, for training, add one round
 * of gradient atomics per block. */
#ifndef GD_INFER_BLOCKS_PER_CU
#define GD_INFER_BLOCKS_PER_CU 3
#endif
#define GD_TRAIN_BLOCKS_PER_CU 1
/* The weight sum of a training step is accumulated in GD_WSUM_N parts
 * (block b adds to part b % GD_WSUM_N: same-address atomics from a
 * thousand blocks serialise), summed by the training kernel. */
#define GD_WSUM_OFF 8
#define GD_WSUM_N 16
#define GD_ACC_FLOATS (GD_WSUM_OFF + GD_WSUM_N)
#define GD_TRAIN_THREADS 256   /* 4 waves: one per SIMD, its gradient sums in the accumulation registers */

/* Training record (16 floats):
 *   0..2 position, 3..5 viewer-side normal, 6..8 outgoing direction,
 *   9..11 sampled direction, 12 BSDF pdf of it, 13 weight, 14 luminance of
 *   the throughput before the bounce, 15 unused.
 * Auxiliary record (float4): radiance of the path when the record was
 * closed (rgb) and the path slot (int bits). */
struct GuideState {
    float* theta;       /* FP32 master parameters */
    _Float16* w16i;     /* FP16 weights: forward tiles, then (w16b) backward tiles */
    _Float16* w16b;
    float* adam_m;
    float* adam_v;
    float* grad;
    float* part;        /* per training block: its gradient sums, GD_NPARAM each */
    int32_t part_blocks;
    float* acc;         /* [0] weight sum, [1] loss sum, [2] records trained, [GD_WSUM_OFF ..] weight-sum parts */
    float* keep;        /* probability that a render keeps a path vertex as a training record */
    int32_t* rec_count;
    unsigned long long* trained;  /* running total of records over all Adam steps */
    unsigned long long* inferred; /* running total of path vertices evaluated by k_guide_infer */
    float* rec;         /* capacity x GD_REC */
    float4* rec_aux;    /* capacity */
    int32_t* open_rec;  /* per path slot: index of its open record or -1 */
    _Float16* out;      /* per queue position of the bounce: GD_PARAMS head outputs */
    float4* in_a;       /* per queue position: position in [0,1]^3, normal.xy (FP16 pair) (k_guide_gather) */
    float2* in_b;       /* per queue position: (normal.z, wo.x), (wo.y, wo.z) as FP16 pairs */
    int32_t rec_cap;
    float lo[3], inv_ext[3];
};

/* Output of W4 that inference lane group g keeps in slot s (0..11), or -1. */
/* Head outputs of a path vertex as the network kernel stores them (48
 * halves, 96 B): the 12 halves of lane group q hold lobes 2q and 2q + 1 (5
 * halves each: mean direction, concentration, mixture weight), group 0 also
 * the probability of sampling the guide at half GD_ALPHA_OFF; the others
 * are zero.  Each lane of the network kernel writes its group as one
 * contiguous 24-byte piece, so every 96-byte row is written whole. */
#define GD_ALPHA_OFF 10
__host__ __device__ inline int gd_lobe_off(int j) { return 12 * (j >> 1) + 5 * (j & 1); }

__host__ __device__ inline int gd_head_output(int g, int s)
{
    return s < 10 ? 10 * g + s : (s == 10 && g == 0 ? 40 : -1);
}

/* Position of master parameter i in the inference copy (-1: biases, and the
 * unused outputs 41..47 of W4). */
__host__ __device__ inline int gd_w16i_slot(int i)
{
    int off, I, base, KB, o, c;
    if (i < GD_OFF_B1) { off = GD_OFF_W1; I = 80; base = GD_W16I_W1; KB = 3; }
    else if (i >= GD_OFF_W2 && i < GD_OFF_B2) { off = GD_OFF_W2; I = 64; base = GD_W16I_W2; KB = 2; }
    else if (i >= GD_OFF_W3 && i < GD_OFF_B3) { off = GD_OFF_W3; I = 64; base = GD_W16I_W3; KB = 2; }
    else if (i >= GD_OFF_W4 && i < GD_OFF_B4) { off = GD_OFF_W4; I = 64; base = GD_W16I_W4; KB = 2; }
    else return -1;
    o = (i - off) / I;
    c = (i - off) % I;
    int ot, m;
    if (off == GD_OFF_W4) {
        int g, s;
        if (o < 40) { g = o / 10; s = o % 10; }
        else if (o == 40) { g = 0; s = 10; }
        else return -1;
        ot = s >> 2;
        m = 4 * g + (s & 3);
    } else {
        ot = o >> 4;
        m = o & 15;
    }
    const int kb = c >> 5, r = c & 31;
    int q, j;
    if (off == GD_OFF_W1) {
        q = r >> 3;
        j = r & 7;
    } else if (r < 16) {
        q = r >> 2;
        j = r & 3;
    } else {
        q = (r - 16) >> 2;
        j = 4 + ((r - 16) & 3);
    }
    return base + (((ot * KB + kb) * 64 + m + 16 * q) << 3) + j;
}

/* Lane group and slot (0..11) in which the head holds output o, or false. */
__host__ __device__ inline bool gd_head_slot(int o, int* g, int* s)
{
    if (o < 40) { *g = o / 10; *s = o % 10; return true; }
    if (o == 40) { *g = 0; *s = 10; return true; }
    return false;
}

/* Position of master parameter i in the backward tiles (-1: W1, biases, and
 * the unused outputs 41..47 of W4). */
__host__ __device__ inline int gd_w16b_slot(int i)
{
    int off, base;
    if (i >= GD_OFF_W2 && i < GD_OFF_B2) { off = GD_OFF_W2; base = GD_W16B_W2; }
    else if (i >= GD_OFF_W3 && i < GD_OFF_B3) { off = GD_OFF_W3; base = GD_W16B_W3; }
    else if (i >= GD_OFF_W4 && i < GD_OFF_B4) { off = GD_OFF_W4; base = GD_W16B_W4; }
    else return -1;
    const int o = (i - off) / GD_H, c = (i - off) % GD_H;  /* output row, input column */
    const int it = c >> 4, m = c & 15;
    int kb, q, j;
    if (off == GD_OFF_W4) {
        int g, s;
        if (!gd_head_slot(o, &g, &s)) return -1;
        kb = s >> 3;
        q = g;
        j = s & 7;
    } else {
        kb = o >> 5;
        const int r = o & 31;
        q = (r & 15) >> 2;
        j = (r & 3) + (r >= 16 ? 4 : 0);
    }
    return base + (((it * 2 + kb) * 64 + m + 16 * q) << 3) + j;
}

/* Whether training computes a gradient for master parameter i (all but the
 * unused outputs 41..47 of W4 and their biases). */
__host__ __device__ inline bool gd_trained(int i)
{
    if (i >= GD_OFF_W4 && i < GD_OFF_B4) return (i - GD_OFF_W4) / GD_H <= 40;
    if (i >= GD_OFF_B4) return i - GD_OFF_B4 <= 40;
    return true;
}

struct GuideHost;  /* forward declaration for the ABI layer */

hipError_t launch_guide_infer(const struct HostDevScene& s, const struct PathState& ps, const GuideState& g,
                              int qsel, int grid, hipStream_t st);
hipError_t launch_guide_forward(const GuideState& g, const float* rec, int n, float* y, int grid, hipStream_t st);
hipError_t launch_guide_head(const GuideState& g, const float* rec, int n, _Float16* out, int grid, hipStream_t st);
/* Training works on chunks of records [off, off + max_n) of the stored ones. */
hipError_t launch_guide_weights(const GuideState& g, const struct PathState& ps, int off, int max_n, int grid,
                                hipStream_t st);
hipError_t launch_guide_zero_acc(const GuideState& g, hipStream_t st);
hipError_t launch_guide_train(const GuideState& g, const float* rec, const int32_t* count, int off, int max_n,
                              int grid, hipStream_t st);
hipError_t launch_guide_adam(const GuideState& g, float lr, int step, const int32_t* count, int off, int max_n,
                             hipStream_t st);
hipError_t launch_guide_reset_records(const GuideState& g, int slots, hipStream_t st);
/* scale = 0, after a training pass: the keep probability for the next pass
 * from the vertices this one offered (records counted / keep), aiming at 90%
 * of the record capacity.  scale > 0: the probability times scale (passes of
 * another size), at most 1. */
hipError_t launch_guide_keep(const GuideState& g, float scale, hipStream_t st);

#endif
