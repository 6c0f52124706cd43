// Learned PPE warm start on gfx950 MFMA.
//
// Network: edge MLP e_ij -> 64 -> 64 (ReLU), sum over the neighbour list
// scaled by 1/64, node MLP 64 -> 64 (ReLU) -> 1.
//
// The forward kernel (k_mlp_forward) gives each wave 16 consecutive
// particles and walks their neighbour lists slot by slot, with every layer
// computed transposed so that the 16 MFMA columns are the 16 particles; the
// per-particle sums then stay in registers (layout at the kernel).
//
// The backward kernel owns one particle per wave64 (grid-stride over
// particles) and walks its neighbour list in tiles of 16 edges:
//   layer 1  D[16 x 64] = E[16 x 32] W1^T[32 x 64]   4 x v_mfma_f32_16x16x32_bf16
//            (the 6 edge inputs padded to K = 32 with zeros)
//   layer 2  D[16 x 64] = H1[16 x 64] W2^T[64 x 64]  8 x v_mfma_f32_16x16x32_bf16
// MFMA operand layout (16x16x32): lane l holds row (l mod 16) of A and
// column (l mod 16) of B, K elements 8*(l/16) .. 8*(l/16)+7; the FP32
// result holds rows 4*(l/16) .. +3 of column (l mod 16).  The layer-1
// result is re-laid out into the layer-2 A operand through a 16 x 64 BF16
// LDS tile.  The node MLP (64 x 64 per particle, 1/80 of the edge work)
// runs on the VALU with the same BF16 weights.
//
// Backward recomputes the forward per particle (two passes over its edges)
// and accumulates weight gradients with v_mfma_f32_16x16x16_bf16, whose A
// and B operands are exactly the FP32 result layout above (K = the 16 edges
// of a tile), so dW2 += dZ2^T H1 and dW1 += dZ1^T E need no re-layout.
// dH1 = dZ2 W2 goes through a second LDS tile.  Each persistent block
// writes its partial gradient; a second kernel sums the partials in a fixed
// order (deterministic), then Adam updates the FP32 master copy and the
// BF16 copy is refreshed.
#include "learned.hpp"

namespace sophia {

namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// Particles of this block: eight contiguous chunks, one per XCD (workgroups
// are dealt to XCDs round robin), so the particles an XCD works on at one
// time are consecutive and share their neighbour-list lines and neighbour
// fields in that XCD's L2.  gridDim.x is a multiple of 8.
struct PartIter {
  int i, end, stride;
};
__device__ inline PartIter xcd_particles(int n) {
  const int x = blockIdx.x & 7, b = blockIdx.x >> 3, per = gridDim.x >> 3;
  const int chunk = (n + 7) >> 3;
  const int lo = x * chunk;
  return PartIter{lo + b, min(n, lo + chunk), per};
}

__device__ inline f4 mfma32(bf8 a, bf8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
__device__ inline f4 mfma16(bf4 a, bf4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

struct Net {
  const __bf16* wbf;
  const float* master;  // biases are read in FP32
};

// B operands of layer 1 (B[k][n] = W1[n][k], k < 6) and layer 2
// (B[k][n] = W2[n][k]) for this lane
__device__ inline void load_fwd_operands(const Net& net, int lane, bf8 (&B1)[4], bf8 (&B2)[4][2]) {
  const int kq = 8 * (lane >> 4), nl = lane & 15;
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    const int nn = nl + 16 * nb;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int k = kq + j;
      B1[nb][j] = k < kEin ? net.wbf[OFF_W1 + nn * kEin + k] : (__bf16)0.0f;
      B2[nb][0][j] = net.wbf[OFF_W2 + nn * kH + k];
      B2[nb][1][j] = net.wbf[OFF_W2 + nn * kH + 32 + k];
    }
  }
}

template <typename R>
struct EdgeTile {
  bf8 a1;            // layer-1 A operand (lanes 0-15 carry the 6 inputs)
  uint32_t valid;    // bit e: edge e of the tile exists
};

// inputs of edge (i, j): (x_i - x_j)/h_i, |x_i - x_j|/h_i, rho_j/rho_f, s_j
template <typename R>
__device__ inline EdgeTile<R> edge_tile(const SphView<R>& v, const float2* nf, int i, int c, int cnt, int lane,
                                        float ein[kEin]) {
  EdgeTile<R> t;
  for (int q = 0; q < 8; q++) t.a1[q] = (__bf16)0.0f;
  for (int q = 0; q < kEin; q++) ein[q] = 0.f;
  bool ok = false;
  if (lane < 16) {
    const int e = c * 16 + lane;
    if (e < cnt) {
      const uint32_t ent = v.nbr[nbr_slot(e, i, v.n)];
      if (!(ent & kNbrFlag)) {
        const int j = (int)ent;
        const double ih = 1.0 / v.h[i];
        double ex, ey, ez;
        pdiff(v, v.x[i], v.y[i], v.z[i], j, ex, ey, ez);
        const double dx = ex * ih, dy = ey * ih, dz = ez * ih;
        const float2 f = nf[j];
        ein[0] = (float)dx;
        ein[1] = (float)dy;
        ein[2] = (float)dz;
        ein[3] = (float)sqrt(dx * dx + dy * dy + dz * dz);
        ein[4] = f.x;
        ein[5] = f.y;
        ok = true;
      }
    }
    for (int q = 0; q < kEin; q++) t.a1[q] = (__bf16)ein[q];
  }
  t.valid = (uint32_t)(__ballot(ok) & 0xffffull);
  return t;
}

// layer 1 + ReLU for one tile; z1 in result layout, H (LDS) gets BF16 h1
__device__ inline void layer1(const Net& net, const bf8 (&B1)[4], bf8 a1, int lane, f4 (&h1)[4], __bf16 (*H)[kH + 8]) {
  const int nl = lane & 15, er = 4 * (lane >> 4);
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    f4 z = {0.f, 0.f, 0.f, 0.f};
    z = mfma32(a1, B1[nb], z);
    const float b = net.master[OFF_B1 + nl + 16 * nb];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      float hv = fmaxf(z[r] + b, 0.f);
      __bf16 hb = (__bf16)hv;
      h1[nb][r] = (float)hb;
      H[er + r][nl + 16 * nb] = hb;
    }
  }
}

// layer 2 pre-activation for one tile (A operand read back from H)
__device__ inline void layer2(const bf8 (&B2)[4][2], int lane, __bf16 (*H)[kH + 8], f4 (&z2)[4]) {
  bf8 a0, a1;
  const int e = lane & 15, kq = 8 * (lane >> 4);
#pragma unroll
  for (int j = 0; j < 8; j++) {
    a0[j] = H[e][kq + j];
    a1[j] = H[e][32 + kq + j];
  }
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    f4 z = {0.f, 0.f, 0.f, 0.f};
    z = mfma32(a0, B2[nb][0], z);
    z = mfma32(a1, B2[nb][1], z);
    z2[nb] = z;
  }
}

// node MLP forward: z3 = W3 m + b3 (lane l = output l), y = W4 relu(z3) + b4
__device__ inline float node_forward(const Net& net, const float* Ms, const __bf16* W3s, int lane, float& z3) {
  float z = net.master[OFF_B3 + lane];
  for (int k = 0; k < kH; k++) z += (float)W3s[lane * kH + k] * Ms[k];
  z3 = z;
  float g = fmaxf(z, 0.f);
  float y = (float)net.wbf[OFF_W4 + lane] * g;
  for (int off = 32; off > 0; off >>= 1) y += __shfl_xor(y, off);
  return y + net.master[OFF_B4];
}

// Forward over the edges of particle i: returns the aggregated message
// (ready in Ms, LDS) -- shared by the forward and backward kernels.
template <typename R>
__device__ inline void aggregate(const Net& net, const SphView<R>& v, const float2* nf, int i, int lane,
                                 const bf8 (&B1)[4], const bf8 (&B2)[4][2], __bf16 (*H)[kH + 8], float* Ms) {
  const int cnt = min(v.nbr_cnt[i], v.nbr_cap);
  const int nl = lane & 15, er = 4 * (lane >> 4);
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  float ein[kEin];
  for (int c = 0; c * 16 < cnt; c++) {
    EdgeTile<R> t = edge_tile(v, nf, i, c, cnt, lane, ein);
    f4 h1[4];
    layer1(net, B1, t.a1, lane, h1, H);
    __syncthreads();
    f4 z2[4];
    layer2(B2, lane, H, z2);
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      const float b = net.master[OFF_B2 + nl + 16 * nb];
#pragma unroll
      for (int r = 0; r < 4; r++)
        if ((t.valid >> (er + r)) & 1u) acc[nb] += fmaxf(z2[nb][r] + b, 0.f);
    }
    __syncthreads();
  }
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    acc[nb] += __shfl_xor(acc[nb], 16);
    acc[nb] += __shfl_xor(acc[nb], 32);
  }
  if (lane < 16) {
#pragma unroll
    for (int nb = 0; nb < 4; nb++) Ms[lane + 16 * nb] = acc[nb] * (1.0f / 64.0f);
  }
  __syncthreads();
}

// Forward kernel, weights as the A operand.  Each layer is computed
// transposed, Z^T[hidden x 16 edges] = W[hidden x K] E^T[K x 16 edges], so a
// layer's MFMA result (lane l: hidden units 16 nb + 4 (l / 16) + r of edge
// l mod 16) is already the next layer's B operand once the hidden dimension
// is enumerated in that order: K-step s of layer 2 takes, in lane l,
// hidden-1 units 16 (2 s + j / 4) + 4 (l / 16) + j mod 4 (j = 0..7), and the
// W2 A operand is loaded in the same order.  No LDS transpose between the
// layers.  Edge inputs are gathered 64 at a time (one edge per lane) into an
// LDS tile; the messages are summed per lane over the tiles and then over
// the 16 lanes of each group.  Biases are added in FP32 after each MFMA, as
// in the backward kernel's recomputation.
struct FwdOperands {
  bf4 A1[4];      // W1 rows 16 nb + (l mod 16), inputs 4 (l / 16) .. +3 (16x16x16 layout): K = 0..5 the
                  // edge inputs, K = 6, 7 the BF16 high and low parts of b1 (inputs 6, 7 are 1), rest 0
  bf8 A2[4][2];   // W2 rows 16 mb + (l mod 16), hidden-1 units in the layer-1 result order
  f4 C2[4];       // b2 in the layer-2 result layout (units 16 mb + 4 (l / 16) + r), the accumulator start
};

__device__ inline void load_fwd_t(const Net& net, int lane, FwdOperands& w) {
  const int m = lane & 15, g = lane >> 4;
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int k = 4 * g + j;
      const float b1 = net.master[OFF_B1 + 16 * nb + m];
      const __bf16 b1h = (__bf16)b1;
      w.A1[nb][j] = k < kEin ? net.wbf[OFF_W1 + (16 * nb + m) * kEin + k]
                  : k == kEin ? b1h : k == kEin + 1 ? (__bf16)(b1 - (float)b1h) : (__bf16)0.0f;
    }
  }
#pragma unroll
  for (int mb = 0; mb < 4; mb++)
#pragma unroll
    for (int sk = 0; sk < 2; sk++)
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int u = 16 * (2 * sk + j / 4) + 4 * g + (j & 3);
        w.A2[mb][sk][j] = net.wbf[OFF_W2 + (16 * mb + m) * kH + u];
      }
#pragma unroll
  for (int mb = 0; mb < 4; mb++)
#pragma unroll
    for (int r = 0; r < 4; r++) w.C2[mb][r] = net.master[OFF_B2 + 16 * mb + 4 * g + r];
}

// ReLU through the sign bit: an integer max with 0 maps every negative
// float (and -0) to +0 and leaves the rest unchanged, without the NaN
// canonicalisation a float max needs
__device__ inline float relu_bits(float x) { return __int_as_float(max(__float_as_int(x), 0)); }
// ReLU of two BF16 values packed in one dword (integer max per half)
__device__ inline uint32_t relu_pk_bf16(float a, float b) {
  typedef short s2 __attribute__((ext_vector_type(2)));
  typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
  typedef float fl2 __attribute__((ext_vector_type(2)));
  const fl2 f = {a, b};
  const s2 v = __builtin_bit_cast(s2, __builtin_convertvector(f, bf2));
  const s2 z = {0, 0};
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(v, z));
}

// Lanes of one wave exchange the edge inputs through LDS.  A wave's LDS
// instructions execute in issue order, so ordering the compiler's code
// motion with wavefront-scope fences is all the synchronisation needed.
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kFwdWaves = 4;    // waves per block of the forward kernel
constexpr int kFwdGrid = 1024;  // blocks at most: 4 waves per SIMD on 256 CUs

// Groups of 16 particles of this wave: eight contiguous chunks of groups,
// one per XCD (blocks are dealt to XCDs round robin), so the groups an XCD
// works on at one time are consecutive and share neighbour fields in that
// XCD's L2.  gridDim.x is a multiple of 8.
struct GroupIter {
  int g, end, stride;
};
__device__ inline GroupIter xcd_groups(int ng) {
  const int x = blockIdx.x & 7, per = gridDim.x >> 3;
  const int chunk = (ng + 7) >> 3;
  const int first = x * chunk + (blockIdx.x >> 3) * kFwdWaves + (int)(threadIdx.x >> 6);
  return GroupIter{first, min(ng, (x + 1) * chunk), per * kFwdWaves};
}

// Forward kernel.  One wave owns a group of 16 consecutive particles and
// walks their neighbour lists slot by slot, four slots per gather round
// (lane l gathers the edge of particle l mod 16 at slot k0 + l / 16).  Each
// layer is computed transposed, Z^T[hidden x 16] = W[hidden x K] E^T[K x 16],
// with the 16 MFMA columns being the group's 16 particles at one slot:
//   - a layer's result (lane l: hidden units 16 nb + 4 (l / 16) + r of
//     particle l mod 16) is the next layer's B operand once the hidden
//     dimension is enumerated in that order (K-step s of the next layer
//     takes, in lane l, units 16 (2 s + j / 4) + 4 (l / 16) + j mod 4,
//     j = 0..7; the A operand holds the weights in the same order);
//   - the ReLU messages are summed over the slots in registers, so the
//     aggregated message of particle l mod 16 ends up spread over the four
//     lanes l mod 16 + 16 q with no cross-lane reduction, already in the
//     node layer's B-operand order.
// The node layer runs on MFMA as well: the FP32 message is split into BF16
// high and low parts (hi + lo equals it to 2^-16 relative), so the result
// matches an FP32 message against the BF16 weights W3.  Each edge carries
// two extra inputs equal to 1 (0 for a slot past the particle's list), and
// b1 enters layer 1 as its BF16 high and low parts against them; a slot with
// no edge then has h1 = 0 exactly.  Layer 2 starts its accumulation from b2
// (FP32), so an absent edge adds exactly ReLU(b2), which is taken off once
// per particle at the end: no per-slot masking.  b3 is added in FP32.  Edge
// inputs use the pos4 position records.
template <typename R>
__global__ __launch_bounds__(64 * kFwdWaves, 4) void k_mlp_forward(Net net, SphView<R> v,
                                                                   const float2* __restrict__ nf,
                                                                   const int* __restrict__ row,
                                                                   const double* __restrict__ pprev, double scale,
                                                                   double* __restrict__ guess) {
  __shared__ bf8 A3s[4][2][64];                         // W3 as the node layer's A operand, per lane
  __shared__ __attribute__((aligned(16))) float bs[2 * kH];  // b3, W4 (W4 from its BF16 copy)
  __shared__ bf8 Et[kFwdWaves][64];                     // edge inputs of one gather round, per wave
  const int tid = threadIdx.x, lane = tid & 63;
  const int m = lane & 15, g4 = lane >> 4;
  for (int q = tid; q < 4 * 2 * 64; q += blockDim.x) {
    const int mb = q >> 7, sk = (q >> 6) & 1, l = q & 63;
    bf8 a;
#pragma unroll
    for (int j = 0; j < 8; j++)
      a[j] = net.wbf[OFF_W3 + (16 * mb + (l & 15)) * kH + 16 * (2 * sk + j / 4) + 4 * (l >> 4) + (j & 3)];
    A3s[mb][sk][l] = a;
  }
  if (tid < kH) {
    bs[tid] = net.master[OFF_B3 + tid];
    bs[kH + tid] = (float)net.wbf[OFF_W4 + tid];
  }
  FwdOperands w;
  load_fwd_t(net, lane, w);
  __syncthreads();
  bf8* E = Et[tid >> 6];
  const bf4* E4 = reinterpret_cast<const bf4*>(E);  // inputs 0-3 and 4-7 of each edge
  const float b4 = net.master[OFF_B4];
  const int ng = (v.n + 15) >> 4;
  for (GroupIter it = xcd_groups(ng); it.g < it.end; it.g += it.stride) {
    const int i = it.g * 16 + m;
    const bool live = i < v.n;
    const bool isrow = live && row[i] != 0;
    const int cnt = isrow ? min(v.nbr_cnt[i], v.nbr_cap) : 0;
    int maxc = cnt;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) maxc = max(maxc, __shfl_xor(maxc, off));
    float y = 0.f;
    if (__ballot(isrow) != 0) {  // wave-uniform
      const float4 pi = isrow ? v.pos4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
      const float ih = isrow ? (float)(1.0 / (double)v.h[i]) : 0.f;
      float acc[4][4] = {};
      int nvalid = 0;  // edges this lane gathered (particle l mod 16)
      uint32_t nxt = g4 < cnt ? v.nbr[nbr_slot(g4, i, v.n)] : kNbrFlag;
      for (int k0 = 0; k0 < maxc; k0 += 4) {
        const uint32_t ent = nxt;
        const int kn = k0 + 4 + g4;
        if (k0 + 4 < maxc) nxt = kn < cnt ? v.nbr[nbr_slot(kn, i, v.n)] : kNbrFlag;
        // an edge outside the list gets all-zero inputs, bias inputs
        // included, so its layer-1 and layer-2 sums are exactly 0 and it adds
        // nothing to the message
        bf8 e8;
#pragma unroll
        for (int q = 0; q < 8; q++) e8[q] = (__bf16)0.0f;
        if (!(ent & kNbrFlag)) {
          nvalid++;
          const int j = (int)ent;
          const float4 pj = v.pos4[j];
          const float2 f = nf[j];
          const float3 d = rel_diff(pi, pj, v.pgrid, v.pmask);
          const float dx = d.x * ih, dy = d.y * ih, dz = d.z * ih;
          e8[0] = (__bf16)dx;
          e8[1] = (__bf16)dy;
          e8[2] = (__bf16)dz;
          e8[3] = (__bf16)__builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz);
          e8[4] = (__bf16)f.x;
          e8[5] = (__bf16)f.y;
          e8[6] = (__bf16)1.0f;  // bias inputs
          e8[7] = (__bf16)1.0f;
        }
        wave_lds_sync();  // the previous round's reads of E are issued
        E[lane] = e8;
        wave_lds_sync();
#pragma unroll 2
        for (int t = 0; t < 4; t++) {
          bf4 b;
#pragma unroll
          for (int q = 0; q < 4; q++) b[q] = (__bf16)0.0f;
          if (g4 < 2) b = E4[2 * (16 * t + m) + g4];
          // layer 1 (16 hidden x 16 particles per block nb, K = 16, b1
          // included), ReLU, packed straight into the layer-2 B operand
          uint32_t B2u[2][4];
#pragma unroll
          for (int nb = 0; nb < 4; nb++) {
            f4 z = {0.f, 0.f, 0.f, 0.f};
            z = mfma16(w.A1[nb], b, z);
            B2u[nb >> 1][2 * (nb & 1) + 0] = relu_pk_bf16(z[0], z[1]);
            B2u[nb >> 1][2 * (nb & 1) + 1] = relu_pk_bf16(z[2], z[3]);
          }
          const bf8 B20 = __builtin_bit_cast(bf8, B2u[0]), B21 = __builtin_bit_cast(bf8, B2u[1]);
          // layer 2 from b2, ReLU and the sum over slots; an absent edge
          // (h1 = 0) adds ReLU(b2), taken off after the last slot
#pragma unroll
          for (int mb = 0; mb < 4; mb++) {
            f4 z = mfma32(w.A2[mb][0], B20, w.C2[mb]);
            z = mfma32(w.A2[mb][1], B21, z);
            acc[mb][0] += relu_bits(z[0]);
            acc[mb][1] += relu_bits(z[1]);
            acc[mb][2] += relu_bits(z[2]);
            acc[mb][3] += relu_bits(z[3]);
          }
        }
      }
      {
        // absent edges among the 4 * rounds slots processed per particle
        int nv = nvalid;
        nv += __shfl_xor(nv, 16);
        nv += __shfl_xor(nv, 32);
        const float absent = (float)(((maxc + 3) & ~3) - nv);
#pragma unroll
        for (int mb = 0; mb < 4; mb++)
#pragma unroll
          for (int r = 0; r < 4; r++) acc[mb][r] -= absent * relu_bits(w.C2[mb][r]);
      }
      // node layer: message M = acc / 64 as BF16 hi + lo parts
      bf8 Bh[2], Bl[2];
#pragma unroll
      for (int sk = 0; sk < 2; sk++)
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const float mv = acc[2 * sk + j / 4][j & 3] * (1.0f / 64.0f);
          const __bf16 hb = (__bf16)mv;
          Bh[sk][j] = hb;
          Bl[sk][j] = (__bf16)(mv - (float)hb);
        }
      float yp = 0.f;
#pragma unroll
      for (int mb = 0; mb < 4; mb++) {
        const bf8 a0 = A3s[mb][0][lane], a1 = A3s[mb][1][lane];
        f4 z = {0.f, 0.f, 0.f, 0.f};
        z = mfma32(a0, Bl[0], z);
        z = mfma32(a1, Bl[1], z);
        z = mfma32(a0, Bh[0], z);
        z = mfma32(a1, Bh[1], z);
        const float4 bb = *reinterpret_cast<const float4*>(bs + 16 * mb + 4 * g4);
        const float4 ww = *reinterpret_cast<const float4*>(bs + kH + 16 * mb + 4 * g4);
        yp += ww.x * fmaxf(z[0] + bb.x, 0.f) + ww.y * fmaxf(z[1] + bb.y, 0.f) + ww.z * fmaxf(z[2] + bb.z, 0.f) +
              ww.w * fmaxf(z[3] + bb.w, 0.f);
      }
      yp += __shfl_xor(yp, 16);
      yp += __shfl_xor(yp, 32);
      y = yp + b4;
    }
    if (g4 == 0 && live) guess[i] = isrow ? pprev[i] + scale * (double)y : pprev[i];
  }
}

template <typename R>
__global__ __launch_bounds__(64) void k_mlp_backward(Net net, SphView<R> v, const float2* __restrict__ nf,
                                                      const int* __restrict__ row, const double* __restrict__ pprev,
                                                      const double* __restrict__ target, double inv_scale,
                                                      float inv_rows, float* __restrict__ partials) {
  __shared__ __bf16 H[16][kH + 8];
  __shared__ __bf16 DZ[16][kH + 8];
  __shared__ __bf16 EIN[16][16];
  __shared__ float Ms[kH];
  __shared__ float DG[kH];
  __shared__ float DM[kH];
  __shared__ __bf16 W3s[kH * kH];
  const int lane = threadIdx.x;
  const int nl = lane & 15, er = 4 * (lane >> 4), kq = 8 * (lane >> 4);
  for (int q = lane; q < kH * kH; q += 64) W3s[q] = net.wbf[OFF_W3 + q];
  bf8 B1[4], B2[4][2];
  load_fwd_operands(net, lane, B1, B2);
  // B operand of dH1 = dZ2 W2: B[k = n][col = kout] = W2[n][kout]
  bf8 Bt[4][2];
#pragma unroll
  for (int kb = 0; kb < 4; kb++)
#pragma unroll
    for (int j = 0; j < 8; j++) {
      Bt[kb][0][j] = net.wbf[OFF_W2 + (kq + j) * kH + nl + 16 * kb];
      Bt[kb][1][j] = net.wbf[OFF_W2 + (32 + kq + j) * kH + nl + 16 * kb];
    }
  const float w4 = (float)net.wbf[OFF_W4 + lane];
  __syncthreads();

  f4 gW2[16], gW1[4];
#pragma unroll
  for (int q = 0; q < 16; q++) gW2[q] = (f4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int q = 0; q < 4; q++) gW1[q] = (f4){0.f, 0.f, 0.f, 0.f};
  float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0};
  float gW3[kH];
#pragma unroll
  for (int k = 0; k < kH; k++) gW3[k] = 0.f;
  float gb3 = 0.f, gW4 = 0.f, gb4 = 0.f, loss = 0.f;

  for (PartIter it = xcd_particles(v.n); it.i < it.end; it.i += it.stride) {
    const int i = it.i;
    if (!row[i]) continue;
    // pass 1: forward
    aggregate(net, v, nf, i, lane, B1, B2, H, Ms);
    float z3;
    const float y = node_forward(net, Ms, W3s, lane, z3);
    const float t = (float)((target[i] - pprev[i]) * inv_scale);
    const float err = y - t;
    const float dy = 2.f * err * inv_rows;
    if (lane == 0) {
      loss += err * err * inv_rows;
      gb4 += dy;
    }
    const float g = fmaxf(z3, 0.f);
    gW4 += dy * g;
    const float dg = z3 > 0.f ? dy * w4 : 0.f;
    gb3 += dg;
#pragma unroll
    for (int k = 0; k < kH; k++) gW3[k] += dg * Ms[k];
    DG[lane] = dg;
    __syncthreads();
    {
      float dm = 0.f;
      for (int l = 0; l < kH; l++) dm += (float)W3s[l * kH + lane] * DG[l];
      DM[lane] = dm * (1.0f / 64.0f);
    }
    __syncthreads();
    // pass 2: edge backward
    const int cnt = min(v.nbr_cnt[i], v.nbr_cap);
    float ein[kEin];
    for (int c = 0; c * 16 < cnt; c++) {
      EdgeTile<R> tt = edge_tile(v, nf, i, c, cnt, lane, ein);
      if (lane < 16)
        for (int q = 0; q < 16; q++) EIN[lane][q] = q < kEin ? (__bf16)ein[q] : (__bf16)0.0f;
      f4 h1[4];
      layer1(net, B1, tt.a1, lane, h1, H);
      __syncthreads();
      f4 z2[4];
      layer2(B2, lane, H, z2);
      // dZ2 in result layout; edges outside the list contribute nothing
      f4 dz2[4];
#pragma unroll
      for (int nb = 0; nb < 4; nb++) {
        const float b = net.master[OFF_B2 + nl + 16 * nb];
        const float dmv = DM[nl + 16 * nb];
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const bool on = ((tt.valid >> (er + r)) & 1u) && (z2[nb][r] + b > 0.f);
          const float d = on ? dmv : 0.f;
          dz2[nb][r] = d;
          gb2[nb] += d;
          DZ[er + r][nl + 16 * nb] = (__bf16)d;
        }
      }
      // dW2 tiles: A = dZ2 (rows n, K = edges), B = H1 (K = edges, cols k)
#pragma unroll
      for (int nb = 0; nb < 4; nb++) {
        bf4 a;
        for (int r = 0; r < 4; r++) a[r] = (__bf16)dz2[nb][r];
#pragma unroll
        for (int kb = 0; kb < 4; kb++) {
          bf4 b;
          for (int r = 0; r < 4; r++) b[r] = (__bf16)h1[kb][r];
          gW2[nb * 4 + kb] = mfma16(a, b, gW2[nb * 4 + kb]);
        }
      }
      __syncthreads();
      // dH1 = dZ2 W2 (A from the DZ tile), then dZ1 through the ReLU
      bf8 ad0, ad1;
#pragma unroll
      for (int j = 0; j < 8; j++) {
        ad0[j] = DZ[nl][kq + j];
        ad1[j] = DZ[nl][32 + kq + j];
      }
      bf4 be;
      for (int r = 0; r < 4; r++) be[r] = EIN[er + r][nl];
#pragma unroll
      for (int kb = 0; kb < 4; kb++) {
        f4 dh = {0.f, 0.f, 0.f, 0.f};
        dh = mfma32(ad0, Bt[kb][0], dh);
        dh = mfma32(ad1, Bt[kb][1], dh);
        bf4 a;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const bool on = ((tt.valid >> (er + r)) & 1u) && h1[kb][r] > 0.f;
          const float d = on ? dh[r] : 0.f;
          gb1[kb] += d;
          a[r] = (__bf16)d;
        }
        gW1[kb] = mfma16(a, be, gW1[kb]);
      }
      __syncthreads();
    }
  }
  // partial gradients of this block
  float* P = partials + (size_t)blockIdx.x * (kNParam + 1);
#pragma unroll
  for (int nb = 0; nb < 4; nb++)
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
      const int n = 16 * nb + er + jj;
      if (nl < kEin) P[OFF_W1 + n * kEin + nl] = gW1[nb][jj];
#pragma unroll
      for (int kb = 0; kb < 4; kb++) P[OFF_W2 + n * kH + 16 * kb + nl] = gW2[nb * 4 + kb][jj];
    }
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    float a = gb1[nb], b = gb2[nb];
    a += __shfl_xor(a, 16);
    a += __shfl_xor(a, 32);
    b += __shfl_xor(b, 16);
    b += __shfl_xor(b, 32);
    if (lane < 16) {
      P[OFF_B1 + lane + 16 * nb] = a;
      P[OFF_B2 + lane + 16 * nb] = b;
    }
  }
#pragma unroll
  for (int k = 0; k < kH; k++) P[OFF_W3 + lane * kH + k] = gW3[k];
  P[OFF_B3 + lane] = gb3;
  P[OFF_W4 + lane] = gW4;
  if (lane == 0) {
    P[OFF_B4] = gb4;
    P[kNParam] = loss;
  }
}

__global__ void k_reduce_partials(const float* __restrict__ partials, int nblocks, float* __restrict__ grad,
                                  double* __restrict__ loss) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > kNParam) return;
  double s = 0;
  for (int b = 0; b < nblocks; b++) s += partials[(size_t)b * (kNParam + 1) + p];
  if (p < kNParam) grad[p] = (float)s;
  else *loss = s;
}

// one Adam step; skipped as a whole when the batch loss is not finite, so a
// single bad batch cannot poison the weights
__global__ void k_adam(float* w, const float* g, float* m, float* v, int n, float lr, float b1t, float b2t,
                       const double* loss) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n || !isfinite(*loss)) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float gm = b1 * m[p] + (1.f - b1) * g[p];
  float gv = b2 * v[p] + (1.f - b2) * g[p] * g[p];
  m[p] = gm;
  v[p] = gv;
  float mh = gm / (1.f - b1t), vh = gv / (1.f - b2t);
  w[p] -= lr * mh / (sqrtf(vh) + eps);
}

__global__ void k_to_bf16(const float* w, __bf16* o, int n) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) o[p] = (__bf16)w[p];
}

// sums over rows of r0^2 and diag, and the row count: per-block partials
// (grid-stride), then one block adds them in a fixed order -> scal[0..2]
constexpr int kFeatBlocks = 256;
__device__ inline double block_sum256(double a, double* sh) {
  for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
  __syncthreads();
  return sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ __launch_bounds__(256) void k_feature_partials(int n, const int* row, const double* r0, const double* diag,
                                                          double* part) {
  __shared__ double sh[4];
  double a = 0, b = 0, c = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    if (row[i]) {
      a += r0[i] * r0[i];
      b += diag[i];
      c += 1;
    }
  a = block_sum256(a, sh);
  b = block_sum256(b, sh);
  c = block_sum256(c, sh);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x + 0] = a;
    part[3 * blockIdx.x + 1] = b;
    part[3 * blockIdx.x + 2] = c;
  }
}

__global__ __launch_bounds__(256) void k_feature_fin(const double* part, int nb, double* scal) {
  __shared__ double sh[4];
  for (int q = 0; q < 3; q++) {
    double a = 0;
    for (int k = threadIdx.x; k < nb; k += 256) a += part[3 * k + q];
    a = block_sum256(a, sh);
    if (threadIdx.x == 0) scal[q] = a;
  }
}

template <typename R>
__global__ void k_node_features(int n, const int* row, const double* r0, const double* rho, double rho_f,
                                double inv_rms, float2* nf) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  nf[i] = make_float2((float)(rho[i] / rho_f), row[i] ? (float)(r0[i] * inv_rms) : 0.f);
}

__global__ void k_prev(int n, const int* row, const double* pres, double* x) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = row[i] ? pres[i] : 0.0;
}

}  // namespace

int mlp_param_count() { return kNParam; }

int mlp_alloc(Mlp& m, int n, int cap) {
  size_t N = (size_t)(n > 0 ? n : 1);
  if (!m.master) {
    HIP_TRY(hipMalloc(&m.master, kNParam * 4));
    HIP_TRY(hipMalloc(&m.grad, kNParam * 4));
    HIP_TRY(hipMalloc(&m.adam_m, kNParam * 4));
    HIP_TRY(hipMalloc(&m.adam_v, kNParam * 4));
    HIP_TRY(hipMalloc(&m.wbf, kNParam * 2));
    HIP_TRY(hipMalloc(&m.partials, (size_t)kMlpBwdBlocks * (kNParam + 1) * 4));
    HIP_TRY(hipMalloc(&m.scal, 8 * 8));
    HIP_TRY(hipMalloc(&m.fpart, kFeatBlocks * 3 * 8));
    HIP_TRY(hipHostMalloc((void**)&m.host_scal, 8 * 8, hipHostMallocDefault));
  }
  if (m.n != n || !m.nf) {
    if (m.nf) (void)hipFree(m.nf);
    if (m.r0) (void)hipFree(m.r0);
    if (m.guess) (void)hipFree(m.guess);
    if (m.target) (void)hipFree(m.target);
    if (m.pprev) (void)hipFree(m.pprev);
    HIP_TRY(hipMalloc(&m.pprev, N * 8));
    HIP_TRY(hipMalloc(&m.nf, N * sizeof(float2)));
    HIP_TRY(hipMalloc(&m.r0, N * 8));
    HIP_TRY(hipMalloc(&m.guess, N * 8));
    HIP_TRY(hipMalloc(&m.target, N * 8));
  }
  m.n = n;
  m.cap = cap;
  return 0;
}

void mlp_free(Mlp& m) {
  for (void* p : {(void*)m.master, (void*)m.grad, (void*)m.adam_m, (void*)m.adam_v, (void*)m.wbf, (void*)m.partials,
                  (void*)m.scal, (void*)m.fpart, (void*)m.nf, (void*)m.r0, (void*)m.guess, (void*)m.target, (void*)m.pprev})
    if (p) (void)hipFree(p);
  if (m.host_scal) (void)hipHostFree(m.host_scal);
  m = Mlp{};
}

int mlp_set_params(hipStream_t s, Mlp& m, const float* host, int n) {
  if (n != kNParam) {
    set_error("mlp_set_params: %d parameters given, network has %d", n, kNParam);
    return SOPHIA_E_ARG;
  }
  HIP_TRY(hipMemcpyAsync(m.master, host, kNParam * 4, hipMemcpyHostToDevice, s));
  k_to_bf16<<<grid_for(kNParam), kBlock, 0, s>>>(m.master, m.wbf, kNParam);
  HIP_TRY(hipGetLastError());
  int rc = mlp_reset_adam(s, m);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(s));
  m.ready = true;
  return 0;
}

int mlp_get_params(hipStream_t s, const Mlp& m, float* host, int n) {
  if (n != kNParam) {
    set_error("mlp_get_params: buffer of %d, network has %d", n, kNParam);
    return SOPHIA_E_ARG;
  }
  if (!m.ready) {
    set_error("mlp_get_params: parameters not set");
    return SOPHIA_E_STATE;
  }
  HIP_TRY(hipMemcpyAsync(host, m.master, kNParam * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

int mlp_reset_adam(hipStream_t s, Mlp& m) {
  if (!m.adam_m) return 0;
  int rc = fill_zero_bytes(s, m.adam_m, kNParam * 4);
  if (!rc) rc = fill_zero_bytes(s, m.adam_v, kNParam * 4);
  m.adam_t = 0;
  return rc;
}

template <typename R>
int mlp_features(hipStream_t s, const DevParams& P, const SphView<R>& v, Mlp& m, const int* row, const double* diag,
                 double* scale, double* rr_base) {
  if (v.n <= 0) {
    *scale = 1;
    if (rr_base) *rr_base = 0;
    return 0;
  }
  const int nb = std::min(kFeatBlocks, (v.n + 255) / 256);
  k_feature_partials<<<nb, 256, 0, s>>>(v.n, row, m.r0, diag, m.fpart);
  k_feature_fin<<<1, 256, 0, s>>>(m.fpart, nb, m.scal);
  HIP_TRY(hipMemcpyAsync(m.host_scal, m.scal, 3 * 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (rr_base) *rr_base = m.host_scal[0];  // sum of r0^2 over the rows
  const double rows = m.host_scal[2] > 0 ? m.host_scal[2] : 1;
  double rms = sqrt(m.host_scal[0] / rows);
  const double dmean = m.host_scal[1] / rows;
  if (!(rms > 0)) rms = 1e-300;
  *scale = dmean > 0 ? rms / dmean : 0.0;
  k_node_features<R><<<grid_for(v.n), kBlock, 0, s>>>(v.n, row, m.r0, v.rho, P.rho_fluid, 1.0 / rms, m.nf);
  HIP_TRY(hipGetLastError());
  return 0;
}

template <typename R>
int mlp_forward(hipStream_t s, const DevParams& P, const SphView<R>& v, Mlp& m, const int* row, const double* pprev,
                double scale, double* guess) {
  if (!m.ready) {
    set_error("learned warm start: parameters not set");
    return SOPHIA_E_STATE;
  }
  if (v.n <= 0) return 0;
  Net net{m.wbf, m.master};
  const int ng = (v.n + 15) / 16, waves = (ng + 7) / 8 * 8;
  const int g = waves / kFwdWaves < kFwdGrid ? ((waves / kFwdWaves + 7) / 8) * 8 : kFwdGrid;
  k_mlp_forward<R><<<g, 64 * kFwdWaves, 0, s>>>(net, v, m.nf, row, pprev, scale, guess);
  HIP_TRY(hipGetLastError());
  return 0;
}

template <typename R>
int mlp_backward(hipStream_t s, const DevParams& P, const SphView<R>& v, Mlp& m, const int* row, const double* pprev,
                 const double* target, double scale, int n_rows, bool step, double lr, double* loss) {
  if (!m.ready) {
    set_error("learned warm start: parameters not set");
    return SOPHIA_E_STATE;
  }
  if (v.n <= 0 || n_rows <= 0) return 0;
  Net net{m.wbf, m.master};
  const int g = kMlpBwdBlocks;
  const double inv_scale = scale != 0 ? 1.0 / scale : 0.0;
  k_mlp_backward<R><<<g, 64, 0, s>>>(net, v, m.nf, row, pprev, target, inv_scale, 1.0f / (float)n_rows,
                                     m.partials);
  k_reduce_partials<<<grid_for(kNParam + 1), kBlock, 0, s>>>(m.partials, g, m.grad, m.scal + 4);
  HIP_TRY(hipGetLastError());
  if (step) {
    m.adam_t += 1;
    const float b1t = (float)pow(0.9, m.adam_t), b2t = (float)pow(0.999, m.adam_t);
    k_adam<<<grid_for(kNParam), kBlock, 0, s>>>(m.master, m.grad, m.adam_m, m.adam_v, kNParam, (float)lr, b1t, b2t,
                                                 m.scal + 4);
    k_to_bf16<<<grid_for(kNParam), kBlock, 0, s>>>(m.master, m.wbf, kNParam);
    HIP_TRY(hipGetLastError());
  }
  if (loss) {
    HIP_TRY(hipMemcpyAsync(m.host_scal + 4, m.scal + 4, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *loss = m.host_scal[4];
  }
  return 0;
}

int mlp_gather_previous(hipStream_t s, int n, const int* row, const double* pres, double* x) {
  if (n <= 0) return 0;
  k_prev<<<grid_for(n), kBlock, 0, s>>>(n, row, pres, x);
  HIP_TRY(hipGetLastError());
  return 0;
}

#define INST(R)                                                                                                    \
  template int mlp_features<R>(hipStream_t, const DevParams&, const SphView<R>&, Mlp&, const int*, const double*, \
                               double*, double*);                                                                 \
  template int mlp_forward<R>(hipStream_t, const DevParams&, const SphView<R>&, Mlp&, const int*, const double*,  \
                              double, double*);                                                                   \
  template int mlp_backward<R>(hipStream_t, const DevParams&, const SphView<R>&, Mlp&, const int*, const double*, \
                               const double*, double, int, bool, double, double*);
INST(float)
INST(double)

}  // namespace sophia
