// LSD radix sort (reduce-then-scan) and device-wide exclusive scan for gfx950.
//
// Replaces the reference's external sorter (vrdxCmdSortKeyValueIndirect,
// engine.cc:1218-1219; "reduce-then-scan", DETAILS.md:26-30).  Each pass of
// up to 8 key bits runs three kernels:
//   upsweep   : every block counts the digits of its contiguous chunk with
//               one LDS atomic per key into 16 bank-interleaved copies of the
//               histogram (copy = lane % 16), then writes hist[d][block];
//   scan      : one workgroup per digit turns hist into exclusive offsets
//               (skipped for grids of at most 16 blocks, whose downsweep
//               blocks sum the raw counts themselves);
//   downsweep : every block walks its chunk in 2048-key sub-tiles, ranks keys
//               stably (wave multisplit: ballots + LDS counters, then a
//               cross-wave prefix), reorders the sub-tile in LDS and writes
//               runs of equal digits contiguously to global memory.
// The narrow digit of a key width that is not a multiple of 8 goes first
// unless the caller names another pass (depth keys, see sort.h).  No
// block ever waits on another block, so the sort cannot hang on forward
// progress.  Counts are read from device memory (indirect semantics) or given
// by the host; the grid is fixed by the capacity and each block derives its
// chunk from the count.  Values are 32 or 64 bits.
//
// The exclusive scan is two kernels: block sums, then each block adds the
// sums before it and scans its chunk through LDS.
#pragma clang fp contract(off)

#include "common.h"
#include "device.h"
#include "sort.h"

namespace vkgs {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;
#ifndef VKGS_SORT_IPT
#define VKGS_SORT_IPT 8
#endif
constexpr int kIpt = VKGS_SORT_IPT;  // keys per thread per sub-tile
constexpr int kTile = kThreads * kIpt;  // keys per sub-tile
#ifndef VKGS_SORT_MAX_BLOCKS
#define VKGS_SORT_MAX_BLOCKS 1024
#endif
constexpr int kMaxBlocks = VKGS_SORT_MAX_BLOCKS;
// A scan's apply blocks each sum the partials before them, which stays cheap
// at twice the sort's grid and gives the binning scan more waves in flight
// (6M splats: bin 0.063 -> 0.059 ms; the sort itself is fastest at 1024).
constexpr int kMaxScanBlocks = 2048;
constexpr uint32_t kRawHistBlocks = 16;

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
  const int lane = __lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t t = __shfl_up(v, d, kWave);
    if (lane >= d) v += t;
  }
  return v;
}

// Exclusive scan across the workgroup; s_tmp holds one word per wave.
template <int kBlock>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_tmp, uint32_t* total) {
  constexpr int kNw = kBlock / kWave;
  const int lane = __lane_id();
  const int w = threadIdx.x / kWave;
  const uint32_t incl = wave_inclusive_scan(v);
  if (lane == kWave - 1) s_tmp[w] = incl;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kNw; ++i) {
    const uint32_t t = s_tmp[i];
    off += (i < w) ? t : 0u;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return off + incl - v;
}

// Contiguous chunk of block blockIdx.x; chunks are multiples of kTile.
__device__ __forceinline__ void block_range(uint32_t n, uint32_t& begin, uint32_t& end) {
  const uint64_t per = ((uint64_t)n + gridDim.x - 1) / gridDim.x;
  const uint64_t chunk = (per + kTile - 1) / kTile * kTile;
  const uint64_t b = (uint64_t)blockIdx.x * chunk;
  begin = (uint32_t)(b < n ? b : n);
  end = (uint32_t)(b + chunk < n ? b + chunk : n);
}

// A pass of a pair of passes that the keys' largest value makes unneeded:
// every key's bits from skip_shift up are zero, so that pair would keep the
// order (a stable sort on an all-zero digit), and both of its passes return
// at once.  The pass count still has its parity, so the result lands where
// the full sort would put it.
__device__ __forceinline__ bool pass_skipped(const uint32_t* __restrict__ d_key_max, uint32_t skip_shift) {
  return d_key_max && skip_shift < 32 && (*d_key_max >> skip_shift) == 0u;
}

// The keys' maximum for a sort given d_key_max, without atomics: the first
// pass's blocks store their own maxima in block_max (max_mode 1), and block 0
// of the second pass reduces them into *key_max (max_mode 2), before any pass
// that tests it.
enum : uint32_t { kMaxNone = 0, kMaxPerBlock = 1, kMaxReduce = 2 };

__global__ __launch_bounds__(kThreads) void sort_upsweep(const uint32_t* __restrict__ keys,
                                                         const uint32_t* __restrict__ d_count, uint32_t max_n,
                                                         uint32_t shift, uint32_t bits,
                                                         uint32_t* __restrict__ hist,
                                                         const uint32_t* __restrict__ d_key_max,
                                                         uint32_t skip_shift, uint32_t max_mode,
                                                         uint32_t* __restrict__ block_max,
                                                         uint32_t* __restrict__ key_max) {
  if (pass_skipped(d_key_max, skip_shift)) return;
  __shared__ uint32_t s_wmax[kWaves];
  // a histogram needs no ranking: one LDS atomic per key, into one of 16
  // copies of the histogram chosen by lane (lane l counts into copy l % 16),
  // bin d of copy j at word 16 d + j.  Depth keys crowd into few digit
  // values: lanes with the same digit then mostly count into different
  // copies, whose words for that bin sit in different banks.  (With 4 copies
  // per wave the 4 depth passes took 0.186 ms; with 16 interleaved copies
  // 0.168 ms; with 64, 0.193 ms, the clearing and summing costing more.)
  constexpr int kCopies = 16;
  __shared__ uint32_t s_h[256 * kCopies];
  const uint32_t tid = threadIdx.x;
  uint32_t* h = s_h + (tid % kCopies);
  for (uint32_t i = tid; i < 256 * kCopies; i += kThreads) s_h[i] = 0u;
  __syncthreads();
  const uint32_t n = d_count ? min(*d_count, max_n) : max_n;
  uint32_t begin, end;
  block_range(n, begin, end);
  const uint32_t mask = (1u << bits) - 1u;
  uint32_t kmax = 0;
  for (uint32_t base = begin; base < end; base += kTile) {
    uint32_t k[kIpt];
#pragma unroll
    for (int r = 0; r < kIpt; ++r) {
      const uint32_t i = base + r * kThreads + tid;
      k[r] = i < end ? keys[i] : 0u;  // (0 for the padding: no effect on the maximum)
      kmax = max(kmax, k[r]);
    }
#pragma unroll
    for (int r = 0; r < kIpt; ++r)
      if (base + r * kThreads + tid < end) atomicAdd(&h[((k[r] >> shift) & mask) * kCopies], 1u);
  }
  if (max_mode == kMaxReduce && blockIdx.x == 0) {
    kmax = 0;
    for (uint32_t b = tid; b < gridDim.x; b += kThreads) kmax = max(kmax, block_max[b]);
  }
  if (max_mode != kMaxNone) {
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) kmax = max(kmax, (uint32_t)__shfl_xor((int)kmax, d, kWave));
    if (__lane_id() == 0) s_wmax[tid / kWave] = kmax;
  }
  __syncthreads();
  if (max_mode != kMaxNone && tid == 0) {
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) m = max(m, s_wmax[i]);
    if (max_mode == kMaxPerBlock)
      block_max[blockIdx.x] = m;
    else if (blockIdx.x == 0)
      *key_max = m;
  }
  if (tid <= mask) {
    // thread d reads its bin's copies rotated by d, so the lanes' reads fall
    // in distinct banks
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kCopies; ++j) s += s_h[tid * kCopies + ((j + tid) % kCopies)];
    hist[tid * gridDim.x + blockIdx.x] = s;
  }
}

// Exclusive scan of one digit's row hist[d][0..blocks) in place (one workgroup
// per digit); digit_total[d] receives the row sum.
__global__ __launch_bounds__(kThreads) void sort_scan_rows(uint32_t* __restrict__ hist, uint32_t blocks,
                                                           uint32_t* __restrict__ digit_total,
                                                           const uint32_t* __restrict__ d_key_max,
                                                           uint32_t skip_shift) {
  if (pass_skipped(d_key_max, skip_shift)) return;
  __shared__ uint32_t s_tmp[kWaves];
  uint32_t* row = hist + (size_t)blockIdx.x * blocks;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < blocks; base += kThreads * 4) {
    const uint32_t first = base + threadIdx.x * 4;
    uint32_t x[4], s = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      x[j] = first + j < blocks ? row[first + j] : 0u;
      s += x[j];
    }
    uint32_t all;
    uint32_t run = carry + block_exclusive_scan<kThreads>(s, s_tmp, &all);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (first + j < blocks) row[first + j] = run;
      run += x[j];
    }
    carry += all;
  }
  if (threadIdx.x == 0) digit_total[blockIdx.x] = carry;
}

template <int kBits, typename V>
__global__ __launch_bounds__(kThreads) void sort_downsweep(const uint32_t* __restrict__ kin,
                                                           const V* __restrict__ vin,
                                                           uint32_t* __restrict__ kout, V* __restrict__ vout,
                                                           const uint32_t* __restrict__ d_count, uint32_t max_n,
                                                           uint32_t shift, const uint32_t* __restrict__ hist,
                                                           const uint32_t* __restrict__ digit_total,
                                                           uint32_t raw_hist, const uint32_t* __restrict__ d_key_max,
                                                           uint32_t skip_shift) {
  if (pass_skipped(d_key_max, skip_shift)) return;
  constexpr uint32_t bits = kBits;
  __shared__ uint32_t s_base[256];
  __shared__ uint32_t s_wcnt[kWaves][256];
  __shared__ uint32_t s_start[256];
  __shared__ uint32_t s_tmp[kWaves];
  __shared__ uint32_t s_key[kTile];
  __shared__ V s_val[kTile];

  const uint32_t tid = threadIdx.x, w = tid / kWave, lane = tid % kWave;
  const uint32_t n = d_count ? min(*d_count, max_n) : max_n;
  uint32_t begin, end;
  block_range(n, begin, end);
  const uint32_t mask = (1u << bits) - 1u;
  {
    // global offset of digit tid for this block: keys of smaller digits
    // everywhere plus keys of this digit in earlier blocks.  raw_hist: the
    // row scan was skipped (small grids), so each block sums the raw counts.
    uint32_t total = 0, before = 0;
    if (tid <= mask) {
      if (raw_hist) {
        for (uint32_t b = 0; b < gridDim.x; ++b) {
          const uint32_t h = hist[tid * gridDim.x + b];
          total += h;
          before += b < blockIdx.x ? h : 0u;
        }
      } else {
        total = digit_total[tid];
        before = hist[tid * gridDim.x + blockIdx.x];
      }
    }
    uint32_t all;
    const uint32_t smaller = block_exclusive_scan<kThreads>(total, s_tmp, &all);
    s_base[tid] = tid <= mask ? smaller + before : 0u;
  }
#pragma unroll
  for (int i = 0; i < kWaves; ++i) s_wcnt[i][tid] = 0u;
  __syncthreads();

  for (uint32_t base = begin; base < end; base += kTile) {
    const uint32_t cnt = min((uint32_t)kTile, end - base);
    const uint32_t seg = base + w * (kWave * kIpt);
    uint32_t k[kIpt], rank[kIpt];
    V v[kIpt];
#pragma unroll
    for (int r = 0; r < kIpt; ++r) {
      const uint32_t i = seg + r * kWave + lane;
      const bool valid = i < end;
      k[r] = valid ? kin[i] : 0u;
      v[r] = valid ? vin[i] : V(0);
    }
    // stable rank inside this wave's segment, 64 keys at a time.  The wave
    // owns its counter row, so no atomics: every lane reads its digit's count,
    // then the first lane of each digit group writes count + group size.  LDS
    // operations of one wave complete in issue order, so the reads of a row
    // see the writes of the rows before it and none of its own.
#pragma unroll
    for (int r = 0; r < kIpt; ++r) {
      const bool valid = seg + r * kWave + lane < end;
      const uint32_t d = (k[r] >> shift) & mask;
      const uint64_t peers = match_digit<kBits>(d, valid);
      const uint32_t below = mbcnt(peers);
      const uint32_t old = valid ? s_wcnt[w][d] : 0u;
      __builtin_amdgcn_wave_barrier();
      if (valid && below == 0) s_wcnt[w][d] = old + (uint32_t)__popcll(peers);
      __builtin_amdgcn_wave_barrier();
      rank[r] = old + below;
    }
    __syncthreads();
    // per digit: exclusive prefix over waves, then over digits
    uint32_t total_d = 0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) {
      const uint32_t c = s_wcnt[i][tid];
      s_wcnt[i][tid] = total_d;
      total_d += c;
    }
    uint32_t all;
    const uint32_t start = block_exclusive_scan<kThreads>(total_d, s_tmp, &all);
    s_start[tid] = start;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kIpt; ++r) {
      const bool valid = seg + r * kWave + lane < end;
      if (valid) {
        const uint32_t d = (k[r] >> shift) & mask;
        const uint32_t pos = s_start[d] + s_wcnt[w][d] + rank[r];
        if (pos < (uint32_t)kTile) {
          s_key[pos] = k[r];
          s_val[pos] = v[r];
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kIpt; ++j) {
      const uint32_t idx = j * kThreads + tid;
      if (idx < cnt) {
        const uint32_t kk = s_key[idx];
        const uint32_t d = (kk >> shift) & mask;
        const uint32_t dst = s_base[d] + idx - s_start[d];
        if (dst < n) {
          kout[dst] = kk;
          vout[dst] = s_val[idx];
        }
      }
    }
    __syncthreads();
    s_base[tid] += total_d;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) s_wcnt[i][tid] = 0u;
    __syncthreads();
  }
}

// Scan inputs: plain words, or the tile count of the packed rectangle in the
// high word of 64-bit sort values.
struct LoadWords {
  const uint32_t* p;
  __device__ uint32_t operator()(uint32_t i) const { return p[i]; }
};
struct LoadRectCount {
  const uint64_t* v;
  RectBits rb;
  __device__ uint32_t operator()(uint32_t i) const {
    uint32_t c;
    unpack_rect((uint32_t)(v[i] >> 32), rb, &c);
    return c;
  }
};

template <typename Load>
__global__ __launch_bounds__(kThreads) void scan_reduce(Load in, const uint32_t* __restrict__ d_count,
                                                        uint32_t max_n, uint32_t* __restrict__ partial) {
  __shared__ uint32_t s_tmp[kWaves];
  const uint32_t n = d_count ? min(*d_count, max_n) : max_n;
  uint32_t begin, end;
  block_range(n, begin, end);
  uint32_t s = 0;
  for (uint32_t i = begin + threadIdx.x; i < end; i += kThreads) s += in(i);
  uint32_t all;
  block_exclusive_scan<kThreads>(s, s_tmp, &all);
  if (threadIdx.x == 0) partial[blockIdx.x] = all;
}

// partial[] holds the raw per-block sums of scan_reduce: each block adds up
// those of the blocks before it (a few per thread) instead of a separate
// single-workgroup scan launch; the last block also writes *d_total.
template <typename Load>
__global__ __launch_bounds__(kThreads) void scan_apply(Load in, uint32_t* out, const uint32_t* __restrict__ d_count,
                                                       uint32_t max_n, const uint32_t* __restrict__ partial,
                                                       uint32_t* __restrict__ d_total) {
  // the sub-tile passes through LDS so that global loads and stores are
  // lane-strided (coalesced) while each thread scans kIpt consecutive words;
  // index i is stored at i + i / 32, which spreads a thread's run over banks
  __shared__ uint32_t s_tmp[kWaves];
  __shared__ uint32_t s_x[kTile + kTile / 32];
  const uint32_t tid = threadIdx.x;
  const uint32_t n = d_count ? min(*d_count, max_n) : max_n;
  uint32_t begin, end;
  block_range(n, begin, end);
  uint32_t carry;
  {
    uint32_t before = 0;
    for (uint32_t i = tid; i < blockIdx.x; i += kThreads) before += partial[i];
    block_exclusive_scan<kThreads>(before, s_tmp, &carry);
    if (d_total != nullptr && tid == 0 && blockIdx.x == gridDim.x - 1) *d_total = carry + partial[blockIdx.x];
  }
  for (uint32_t base = begin; base < end; base += kTile) {
#pragma unroll
    for (int r = 0; r < kIpt; ++r) {
      const uint32_t i = r * kThreads + tid;
      s_x[i + (i >> 5)] = base + i < end ? in(base + i) : 0u;
    }
    __syncthreads();
    uint32_t x[kIpt];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kIpt; ++j) {
      const uint32_t i = tid * kIpt + j;
      x[j] = s_x[i + (i >> 5)];
      s += x[j];
    }
    uint32_t all;
    uint32_t run = carry + block_exclusive_scan<kThreads>(s, s_tmp, &all);
#pragma unroll
    for (int j = 0; j < kIpt; ++j) {
      const uint32_t i = tid * kIpt + j;
      s_x[i + (i >> 5)] = run;
      run += x[j];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kIpt; ++r) {
      const uint32_t i = r * kThreads + tid;
      if (base + i < end) out[base + i] = s_x[i + (i >> 5)];
    }
    __syncthreads();
    carry += all;
  }
}

template <typename V>
__global__ void copy_pairs(const uint32_t* __restrict__ ks, const V* __restrict__ vs, uint32_t* __restrict__ kd,
                           V* __restrict__ vd, const uint32_t* __restrict__ d_count, uint32_t max_n) {
  const uint32_t n = d_count ? min(*d_count, max_n) : max_n;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    kd[i] = ks[i];
    vd[i] = vs[i];
  }
}

__global__ void fill_u32(uint32_t* __restrict__ dst, uint32_t value, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = value;
}

__global__ void set_words(uint32_t* __restrict__ dst, Words16 words, uint32_t n) {
  if (threadIdx.x < n) dst[threadIdx.x] = words.w[threadIdx.x];
}

__global__ void fold_status_kernel(const uint32_t* __restrict__ status, const uint32_t* __restrict__ count,
                                   uint32_t* __restrict__ sticky, uint32_t* __restrict__ sticky_count) {
  if (threadIdx.x == 0 && *status != 0u) {
    *sticky = 1u;
    *sticky_count = *count > *sticky_count ? *count : *sticky_count;
  }
}

struct WordChunk {
  uint32_t w[kUploadChunkWords];
};

__global__ __launch_bounds__(256) void set_chunk(uint32_t* __restrict__ dst, WordChunk chunk, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = chunk.w[i];
}

}  // namespace

SortPlan sort_plan(uint32_t max_n, uint32_t val_bytes) {
  SortPlan p;
  p.max_n = max_n;
  uint32_t tiles = ceil_div(max_n ? max_n : 1, kTile);
  p.blocks = tiles < (uint32_t)kMaxBlocks ? tiles : (uint32_t)kMaxBlocks;
  size_t off = 0;
  p.off_alt_key = off;
  off = align_up(off + (size_t)max_n * 4, 256);
  p.off_alt_val = off;
  off = align_up(off + (size_t)max_n * val_bytes, 256);
  p.off_hist = off;
  off = align_up(off + (size_t)256 * p.blocks * 4, 256);
  p.off_digit_total = off;
  off = align_up(off + 256 * 4, 256);
  p.off_block_max = off;
  off = align_up(off + (size_t)p.blocks * 4, 256);
  p.bytes = off;
  return p;
}

namespace {

template <typename V>
hipError_t sort_launch(const SortPlan& plan, const uint32_t* d_count, uint32_t* keys, V* vals, uint32_t key_bits,
                       void* storage, hipStream_t stream, bool input_in_alt = false, uint32_t narrow_pass = 0,
                       uint32_t* d_key_max = nullptr, uint32_t first_bits = 0) {
  if (key_bits > 32) key_bits = 32;
  char* base = static_cast<char*>(storage);
  uint32_t* alt_k = reinterpret_cast<uint32_t*>(base + plan.off_alt_key);
  V* alt_v = reinterpret_cast<V*>(base + plan.off_alt_val);
  if (plan.max_n == 0) return hipSuccess;
  if (key_bits == 0) {
    if (input_in_alt) copy_pairs<V><<<1024, 256, 0, stream>>>(alt_k, alt_v, keys, vals, d_count, plan.max_n);
    return hipGetLastError();
  }
  uint32_t* hist = reinterpret_cast<uint32_t*>(base + plan.off_hist);
  uint32_t* digit_total = reinterpret_cast<uint32_t*>(base + plan.off_digit_total);
  uint32_t* src_k = input_in_alt ? alt_k : keys;
  V* src_v = input_in_alt ? alt_v : vals;
  uint32_t* dst_k = input_in_alt ? keys : alt_k;
  V* dst_v = input_in_alt ? vals : alt_v;
  // small grids: each downsweep block sums the raw block counts itself,
  // which is cheaper than a row-scan launch
  const bool raw = plan.blocks <= kRawHistBlocks;
  // the narrow digit (when key_bits is not a multiple of 8) goes in pass
  // narrow_pass, the first by default
  // digit widths: 8 bits each except the narrow pass's remainder, or with
  // first_bits (1-7) the first pass that many and the last the remainder
  uint32_t width[5];
  uint32_t n_pass;
  if (first_bits > 0 && first_bits < 8 && key_bits > first_bits) {
    n_pass = 1 + (key_bits - first_bits + 7) / 8;
    for (uint32_t p = 0; p < n_pass; ++p) width[p] = p == 0 ? first_bits : 8u;
    width[n_pass - 1] = key_bits - first_bits - 8 * (n_pass - 2);
  } else {
    n_pass = (key_bits + 7) / 8;
    const uint32_t narrow_at = narrow_pass < n_pass ? narrow_pass : n_pass - 1;
    for (uint32_t p = 0; p < n_pass; ++p) width[p] = p == narrow_at ? key_bits - 8 * (n_pass - 1) : 8u;
  }
  // with d_key_max and four passes, passes 0 and 1 find the keys' maximum and
  // passes 2 and 3 run only if it has a bit at or above pass 2's shift (as a
  // pair, so the result stays where four passes put it)
  const bool skipping = d_key_max && n_pass == 4;
  uint32_t* block_max = reinterpret_cast<uint32_t*>(base + plan.off_block_max);
  const uint32_t pair_shift = n_pass == 4 ? width[0] + width[1] : 32u;  // shift of pass 2
  for (uint32_t shift = 0, pass = 0; shift < key_bits; ++pass) {
    const uint32_t bits = width[pass];
    const uint32_t skip_shift = skipping && pass >= 2 ? pair_shift : 32u;
    // the first two passes find the keys' maximum that the later pair tests
    const uint32_t max_mode = !skipping ? kMaxNone : pass == 0 ? kMaxPerBlock : pass == 1 ? kMaxReduce : kMaxNone;
    sort_upsweep<<<plan.blocks, kThreads, 0, stream>>>(src_k, d_count, plan.max_n, shift, bits, hist,
                                                       skipping ? d_key_max : nullptr, skip_shift, max_mode,
                                                       block_max, d_key_max);
    if (!raw)
      sort_scan_rows<<<1u << bits, kThreads, 0, stream>>>(hist, plan.blocks, digit_total,
                                                          skipping ? d_key_max : nullptr, skip_shift);
    switch (bits) {
#define VKGS_DOWNSWEEP(B)                                                                          \
  case B:                                                                                          \
    sort_downsweep<B, V><<<plan.blocks, kThreads, 0, stream>>>(src_k, src_v, dst_k, dst_v, d_count, \
                                                               plan.max_n, shift, hist, digit_total, \
                                                               raw ? 1u : 0u, skipping ? d_key_max : nullptr, \
                                                               skip_shift);                        \
    break;
      VKGS_DOWNSWEEP(1)
      VKGS_DOWNSWEEP(2)
      VKGS_DOWNSWEEP(3)
      VKGS_DOWNSWEEP(4)
      VKGS_DOWNSWEEP(5)
      VKGS_DOWNSWEEP(6)
      VKGS_DOWNSWEEP(7)
      VKGS_DOWNSWEEP(8)
#undef VKGS_DOWNSWEEP
    }
    shift += bits;
    uint32_t* tk = src_k;
    V* tv = src_v;
    src_k = dst_k;
    src_v = dst_v;
    dst_k = tk;
    dst_v = tv;
  }
  if (src_k != keys) copy_pairs<V><<<1024, 256, 0, stream>>>(src_k, src_v, keys, vals, d_count, plan.max_n);
  return hipGetLastError();
}

}  // namespace

hipError_t sort_kv_launch(const SortPlan& plan, const uint32_t* d_count, uint32_t* keys, uint32_t* vals,
                          uint32_t key_bits, void* storage, hipStream_t stream, uint32_t narrow_pass,
                          uint32_t* d_key_max, uint32_t first_bits) {
  return sort_launch<uint32_t>(plan, d_count, keys, vals, key_bits, storage, stream, false, narrow_pass, d_key_max,
                               first_bits);
}

uint32_t sort_passes(uint32_t key_bits) { return ((key_bits > 32 ? 32 : key_bits) + 7) / 8; }

void sort_alt_buffers(const SortPlan& plan, void* storage, uint32_t** keys, uint32_t** vals) {
  char* base = static_cast<char*>(storage);
  *keys = reinterpret_cast<uint32_t*>(base + plan.off_alt_key);
  *vals = reinterpret_cast<uint32_t*>(base + plan.off_alt_val);
}

hipError_t sort_kv_from_alt_launch(const SortPlan& plan, const uint32_t* d_count, uint32_t* keys, uint32_t* vals,
                                   uint32_t key_bits, void* storage, hipStream_t stream) {
  return sort_launch<uint32_t>(plan, d_count, keys, vals, key_bits, storage, stream, true);
}

hipError_t sort_kv64_launch(const SortPlan& plan, const uint32_t* d_count, uint32_t* keys, uint64_t* vals,
                            uint32_t key_bits, void* storage, hipStream_t stream, uint32_t narrow_pass,
                            uint32_t* d_key_max, uint32_t first_bits) {
  return sort_launch<uint64_t>(plan, d_count, keys, vals, key_bits, storage, stream, false, narrow_pass, d_key_max,
                               first_bits);
}

ScanPlan scan_plan(uint32_t max_n) {
  ScanPlan p;
  p.max_n = max_n;
  uint32_t tiles = ceil_div(max_n ? max_n : 1, kTile);
  p.blocks = tiles < (uint32_t)kMaxScanBlocks ? tiles : (uint32_t)kMaxScanBlocks;
  p.off_partial = 0;
  p.bytes = align_up((size_t)p.blocks * 4, 256);
  return p;
}

namespace {

template <typename Load>
hipError_t scan_launch(const ScanPlan& plan, Load in, uint32_t* out, const uint32_t* d_count, uint32_t* d_total,
                       void* storage, hipStream_t stream) {
  uint32_t* partial = reinterpret_cast<uint32_t*>(static_cast<char*>(storage) + plan.off_partial);
  scan_reduce<Load><<<plan.blocks, kThreads, 0, stream>>>(in, d_count, plan.max_n, partial);
  scan_apply<Load><<<plan.blocks, kThreads, 0, stream>>>(in, out, d_count, plan.max_n, partial, d_total);
  return hipGetLastError();
}

}  // namespace

hipError_t scan_exclusive_launch(const ScanPlan& plan, const uint32_t* in, uint32_t* out, const uint32_t* d_count,
                                 uint32_t* d_total, void* storage, hipStream_t stream) {
  return scan_launch(plan, LoadWords{in}, out, d_count, d_total, storage, stream);
}

hipError_t scan_rect_counts_launch(const ScanPlan& plan, const uint64_t* val64, const RectBits& rb, uint32_t* out,
                                   const uint32_t* d_count, uint32_t* d_total, void* storage, hipStream_t stream) {
  return scan_launch(plan, LoadRectCount{val64, rb}, out, d_count, d_total, storage, stream);
}

hipError_t fill_u32_launch(uint32_t* dst, uint32_t value, size_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  size_t blocks = (n + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  fill_u32<<<(uint32_t)blocks, 256, 0, stream>>>(dst, value, n);
  return hipGetLastError();
}

hipError_t fold_status_launch(const uint32_t* status, const uint32_t* count, uint32_t* sticky,
                              uint32_t* sticky_count, hipStream_t stream) {
  fold_status_kernel<<<1, 64, 0, stream>>>(status, count, sticky, sticky_count);
  return hipGetLastError();
}

hipError_t upload_launch(void* dst, const void* src, size_t bytes, hipStream_t stream) {
  if (bytes % 4 != 0) return hipErrorInvalidValue;
  const uint32_t* s = static_cast<const uint32_t*>(src);
  uint32_t* d = static_cast<uint32_t*>(dst);
  for (size_t off = 0; off < bytes / 4; off += kUploadChunkWords) {
    const uint32_t n = (uint32_t)(bytes / 4 - off < kUploadChunkWords ? bytes / 4 - off : kUploadChunkWords);
    WordChunk c;
    for (uint32_t i = 0; i < n; ++i) c.w[i] = s[off + i];
    set_chunk<<<1, 256, 0, stream>>>(d + off, c, n);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t set_words_launch(uint32_t* dst, const Words16& words, uint32_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  if (n > 16) return hipErrorInvalidValue;
  set_words<<<1, 64, 0, stream>>>(dst, words, n);
  return hipGetLastError();
}

}  // namespace vkgs
