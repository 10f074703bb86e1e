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
 * FP16 weight copy (27,648 values), pre-tiled as MFMA A operands so that a
 * wave reads each 16x16 operand as 64 contiguous 8-byte words (no LDS bank
 * conflicts): W1..W4 for the forward pass (tile (ot, kb), lane m + 16 q
 * holds W[16 ot + m][16 kb + 4 q .. +3]), then W2..W4 transposed for the
 * backward pass (tile (it, kb), lane m + 16 q holds W[16 kb + 4 q .. +3][16 it + m]).
 * These 16x16x16 tiles serve training (v_mfma_f32_16x16x16_f16).
 *
 * Inference copy (17,408 values, v_mfma_f32_16x16x32_f16 A operands of
 * 16 rows x 32 inputs, 64 lanes x 8 halves per tile, tiles (ot, kb)):
 * lane m + 16 q, slot j holds W[row(ot, m)][32 kb + col(q, j)].
 *   W1 (inputs padded to 96): col = 8 q + j, the feature order itself.
 *   W2..W4: col = 4 q + j for j < 4, 16 + 4 q + (j - 4) for j >= 4, the
 *   order in which a lane holds the previous layer's outputs (its D tiles
 *   2 kb and 2 kb + 1), so a layer's result is the next one's B operand.
 *   W1..W3: row = 16 ot + m.  W4: row m of tile ot is the output that lane
 *   group m / 4 keeps in slot 4 ot + m % 4: slots 0..9 are lobes 2 g and
 *   2 g + 1 of lane group g (outputs 10 g .. 10 g + 9), slot 10 of group 0
 *   is the guide probability logit (output 40); the other slots are zero.
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
#define GD_NW16F 16384  /* forward tiles */
#define GD_NW16 27648   /* forward + transposed tiles */
#define GD_OFF_W1 0
#define GD_OFF_B1 5120
#define GD_OFF_W2 5184
#define GD_OFF_B2 9280
#define GD_OFF_W3 9344
#define GD_OFF_B3 13440
#define GD_OFF_W4 13504
#define GD_OFF_B4 16576
#define GD_W16_W1 0
#define GD_W16_W2 5120
#define GD_W16_W3 9216
#define GD_W16_W4 13312
#define GD_W16T_W2 16384
#define GD_W16T_W3 20480
#define GD_W16T_W4 24576
#define GD_NW16I 17408   /* inference tiles (16x16x32 layout) */
#define GD_W16I_W1 0
#define GD_W16I_W2 6144
#define GD_W16I_W3 10240
#define GD_W16I_W4 14336
#define GD_KLO -3.0f
#define GD_KHI 10.0f
#define GD_REC 16        /* floats per training record */
#define GD_PARAMS 48     /* halves of head output per path vertex */
#define GD_WCLAMP 100.0f /* cap of a normalised sample weight */
#define GD_TRAIN_CHUNKS 4 /* training steps per render pass (record capacity = 4 batches) */
/* Persistent grids of the network kernels: the blocks one CU holds at once
 * (LDS: 45 KB per inference block, 138 KB per training block).  More blocks
 * would only reload the weights into LDS and, for training, add one round
 * of gradient atomics per block. */
#define GD_INFER_BLOCKS_PER_CU 3
#define GD_TRAIN_BLOCKS_PER_CU 1

/* Training record (16 floats):
 *   0..2 position, 3..5 viewer-side normal, 6..8 outgoing direction,
 *   9..11 sampled direction, 12 BSDF pdf of it, 13 weight, 14 luminance of
 *   the throughput before the bounce, 15 unused.
 * Auxiliary record (float4): radiance of the path when the record was
 * closed (rgb) and the path slot (int bits). */
struct GuideState {
    float* theta;       /* FP32 master parameters */
    _Float16* w16;      /* FP16 weights, training tiles */
    _Float16* w16i;     /* FP16 weights, inference tiles */
    float* adam_m;
    float* adam_v;
    float* grad;
    float* acc;         /* [0] weight sum, [1] loss sum, [2] records trained */
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

/* Positions of master parameter i in the FP16 copy: forward tile slot and
 * transposed tile slot (-1 where none: biases, and W1 has no transposed copy). */
__host__ __device__ inline void gd_w16_slots(int i, int* fwd, int* bwd)
{
    *fwd = -1;
    *bwd = -1;
    int off, O, I, bf, bb;
    if (i < GD_OFF_B1) { off = GD_OFF_W1; O = 64; I = 80; bf = GD_W16_W1; bb = -1; }
    else if (i >= GD_OFF_W2 && i < GD_OFF_B2) { off = GD_OFF_W2; O = 64; I = 64; bf = GD_W16_W2; bb = GD_W16T_W2; }
    else if (i >= GD_OFF_W3 && i < GD_OFF_B3) { off = GD_OFF_W3; O = 64; I = 64; bf = GD_W16_W3; bb = GD_W16T_W3; }
    else if (i >= GD_OFF_W4 && i < GD_OFF_B4) { off = GD_OFF_W4; O = 48; I = 64; bf = GD_W16_W4; bb = GD_W16T_W4; }
    else return;
    int r = i - off, o = r / I, c = r % I;
    *fwd = bf + ((((o >> 4) * (I >> 4) + (c >> 4)) * 64 + (o & 15) + 16 * ((c & 15) >> 2)) << 2) + (c & 3);
    if (bb >= 0)
        *bwd = bb + ((((c >> 4) * (O >> 4) + (o >> 4)) * 64 + (c & 15) + 16 * ((o & 15) >> 2)) << 2) + (o & 3);
}

/* Output of W4 that inference lane group g keeps in slot s (0..11), or -1. */
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

#endif
