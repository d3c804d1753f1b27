// gfx950 kernels of the parallel XOR index (xindex.hpp): candidates over
// span groups, the per-chunk walker, emit and the per-chunk event fold.
// Every kernel works on the column cols[0].
#include "launch.hpp"
#include "xindex.hpp"

namespace ebi {

// Stream reader of the index walks: lanes walk a few KiB of their own
// stream, so one block of look-ahead beyond the next keeps a load in flight.
using XWindow = BitWindowT<2>;

// Step 1 over the span groups of the column: group (k, j) covers spans
// [j, j + kSpansPerLane) of chunk k.  Gorilla: spans from the walker's
// frontier on, guessed in the state it reached there.
__global__ void __launch_bounds__(256) k_xcands(XIndexArgs a) {
  const uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= a.n_groups) return;
  const uint2 gr = a.groups[gi];
  const uint32_t k = gr.x;
  const DevChunk c = a.cols[0].chunks[k];
  const uint32_t bits = c.size * 8u, ns = span_count(bits), sf = a.span_first[k];
  uint32_t j0 = gr.y;
  uint64_t hint = 0;
  if (c.codec == kCodecGorilla) {
    const ChunkWalk w = a.walk[k];
    if (w.frontier >= ns) return;
    j0 = max(j0, w.frontier);
    hint = w.cur;
  }
  const uint32_t j1 = min(gr.y + kSpansPerLane, ns);
  if (j0 < j1)
    span_cands_any<XWindow>(c.codec, reinterpret_cast<const uint8_t *>(c.data), bits, ns, j0, j1, hint, a.cand + sf,
                            a.res + sf);
}

// Wave-wide scans of the walker (lanes past `n` contribute zeros).
__device__ __forceinline__ uint64_t wave_incl_xor_u64(uint64_t x) {
  const uint32_t l = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint64_t y = __shfl_up(x, d, kWave);
    if (l >= (uint32_t)d) x ^= y;
  }
  return x;
}
__device__ __forceinline__ int wave_incl_max_i32(int x) {
  const uint32_t l = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int y = __shfl_up(x, d, kWave);
    if (l >= (uint32_t)d) x = max(x, y);
  }
  return x;
}

// Orders a wave's global stores before its lanes' later loads of them.
__device__ __forceinline__ void wave_fence_global() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// chunk_walk (xindex.hpp) by one wave, 64 spans at a time: span i of a
// block is a hit when its candidate equals the true entry, which for every
// span after the block's first is the previous span's exit -- so the hits
// are the lanes before the first one whose candidate differs from its
// predecessor's exit (a ballot), and their first record indexes, values
// and beta* are scans over the lanes.  A miss is parsed exactly by the
// whole wave (the same walk on every lane); Gorilla's walker blocks (each
// lane one block_cands walk over kBlockSpansPerLane spans) follow a miss
// whose state is not the one the next span was guessed in.  Same results
// as chunk_walk.
template <uint32_t kCodec>
__device__ void wave_walk(const uint8_t *data, uint32_t bits, uint32_t ns, ChunkWalk &w, uint64_t *cand,
                          SpanSum *res, SpanPick *pick, uint32_t stop, bool blocks) {
  const uint32_t l = lane_id();
  const uint32_t end = min(ns, stop), chint = (uint32_t)(w.cur >> 32) & 0xffff;
  uint32_t i0 = w.frontier, bend = 0, bhint = 0;
  auto dead_from = [&](uint32_t from) {
    for (uint32_t i = from + l; i < ns; i += kWave) pick[i] = SpanPick{kEntDead, 0, 0, INT32_MAX};
  };
  while (i0 < end) {
    const uint32_t i = i0 + l;
    const bool valid = i < end;
    const uint64_t c = valid ? cand[i] : kEntGuess;
    SpanSum rs = valid ? res[i] : SpanSum{0, 0, 0, -1, 0};
    const uint64_t prev = __shfl_up(rs.exit, 1, kWave);
    const uint64_t ent = l == 0 ? w.cur : prev;
    const bool link = valid && c == ent;
    const uint64_t nolink = __ballot(!link);
    const uint32_t r = nolink ? (uint32_t)__builtin_ctzll(nolink) : (uint32_t)kWave;
    const uint64_t errs = __ballot(l < r && rs.err != 0);
    const uint32_t cut = errs ? (uint32_t)__builtin_ctzll(errs) + 1 : r;  // spans taken from res
    if (cut) {
      const bool in = l < cut;
      const uint32_t cnt = in ? rs.cnt : 0u;
      const uint64_t xr = in ? rs.xr : 0ull;
      const uint32_t gs = wave_incl_scan_u32(cnt);
      const uint64_t xs = wave_incl_xor_u64(xr);
      const int lb = wave_incl_max_i32(in && rs.lbeta >= 0 ? (int)l : -1);
      const int lb_ex = __shfl_up(lb, 1, kWave);
      const int src = l == 0 ? -1 : lb_ex;
      const int32_t bsrc = (int32_t)__shfl((int)rs.lbeta, src < 0 ? 0 : src, kWave);
      if (in) pick[i] = SpanPick{ent, w.v ^ (xs ^ xr), w.g + (gs - cnt), src < 0 ? w.beta : bsrc};
      const int last = (int)cut - 1;
      const int lb_all = __shfl(lb, last, kWave);
      if (lb_all >= 0) w.beta = (int32_t)__shfl((int)rs.lbeta, lb_all, kWave);
      w.g += __shfl(gs, last, kWave);
      w.v ^= __shfl(xs, last, kWave);
      w.cur = __shfl(rs.exit, last, kWave);
      if (errs) {
        dead_from(i0 + cut);
        w.frontier = ns;
        return;
      }
      i0 += cut;
      if (cut == (uint32_t)kWave || i0 >= end) continue;
    }
    // span i0 misses: parse it from its true entry
    const SpanSum s = span_settle<kCodec, XWindow>(data, w.cur, span_end(i0, bits), i0 + 1 == ns, bits);
    if (l == 0) pick[i0] = SpanPick{w.cur, w.v, w.g, w.beta};
    if (s.err) {
      dead_from(i0 + 1);
      w.frontier = ns;
      return;
    }
    w.g += s.cnt;
    w.v ^= s.xr;
    if (s.lbeta >= 0) w.beta = s.lbeta;
    w.cur = s.exit;
    ++i0;
    const uint32_t st = (uint32_t)(w.cur >> 32) & 0xffff;
    if (blocks && i0 < end && st != (i0 < bend ? bhint : chint)) {
      bend = min(end, i0 + (uint32_t)kWave * kBlockSpansPerLane);
      bhint = st;
      const uint32_t j = i0 + l * kBlockSpansPerLane;
      if (j < bend)
        block_cands<kCodec, XWindow>(data, bits, ns, j, min(j + kBlockSpansPerLane, bend), i0, w.cur, cand, res);
      wave_fence_global();
    }
  }
  w.frontier = end;
}

// Step 2, a wave per chunk, up to span `stop`; blocks: Gorilla's walker
// blocks.
__global__ void __launch_bounds__(256) k_xwalk(XIndexArgs a, uint32_t stop, uint32_t blocks) {
  const uint32_t k = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + threadIdx.x / kWave);
  if (k >= a.nc) return;
  const DevChunk c = a.cols[0].chunks[k];
  const uint32_t bits = c.size * 8u, ns = span_count(bits), sf = a.span_first[k];
  ChunkWalk w = a.walk[k];
  if (w.frontier >= ns) return;
  const uint8_t *data = reinterpret_cast<const uint8_t *>(c.data);
  uint64_t *cand = a.cand + sf;
  SpanSum *res = a.res + sf;
  SpanPick *pick = a.pick + sf;
  switch (c.codec) {
    case kCodecGorilla: wave_walk<kCodecGorilla>(data, bits, ns, w, cand, res, pick, stop, blocks != 0); break;
    case kCodecChimp: wave_walk<kCodecChimp>(data, bits, ns, w, cand, res, pick, stop, false); break;
    case kCodecChimp128: wave_walk<kCodecChimp128>(data, bits, ns, w, cand, res, pick, stop, false); break;
    case kCodecElf: wave_walk<kCodecElf>(data, bits, ns, w, cand, res, pick, stop, false); break;
    default: wave_walk<kCodecElfOnChimp>(data, bits, ns, w, cand, res, pick, stop, false); break;
  }
  if (lane_id() == 0) a.walk[k] = w;
}

// Step 3 over the span groups: checkpoints and the chunk's first event.
__global__ void __launch_bounds__(256) k_xemit(XIndexArgs a) {
  const uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= a.n_groups) return;
  const uint2 gr = a.groups[gi];
  const uint32_t k = gr.x;
  const DevChunk c = a.cols[0].chunks[k];
  const uint32_t bits = c.size * 8u, ns = span_count(bits), sf = a.span_first[k];
  const uint8_t *data = reinterpret_cast<const uint8_t *>(c.data);
  XorCkpt *ck = reinterpret_cast<XorCkpt *>(c.xidx);
  uint64_t ev = ~0ull;
  for (uint32_t i = gr.y; i < min(gr.y + kSpansPerLane, ns); ++i) {
    const SpanPick p = a.pick[sf + i];
    const uint64_t e =
        span_emit_any<XWindow>(c.codec, data, p.ent, p.g0, p.v0, p.beta0, span_end(i, bits), i + 1 == ns, bits, c.n, ck);
    ev = min(ev, e);
  }
  if (ev != ~0ull) atomicMin(reinterpret_cast<unsigned long long *>(a.ev + k), (unsigned long long)ev);
}

// Per chunk: its first event as the serial index pass reports it (a NaN
// end shortens the chunk, an error marks it).
__global__ void __launch_bounds__(256) k_xresolve(XIndexArgs a) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.nc) return;
  const uint64_t ev = a.ev[k];
  if (ev == ~0ull) return;
  const uint32_t code = (uint32_t)(ev & 0xff), rec = (uint32_t)(ev >> 8);
  const ColView &cv = a.cols[0];
  if (code == kNanEnd) {
    if (cv.chunk_end) cv.chunk_end[k] = rec;
    return;
  }
  atomicOr(a.err, code);
  if (cv.chunk_err) atomicOr(cv.chunk_err + k, code);
}

// Initial walker state of every chunk (chunk start, frontier 0), no
// candidate, no event.
__global__ void __launch_bounds__(256) k_xinit(XIndexArgs a, uint32_t n_spans) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_spans) a.cand[i] = kEntGuess;
  if (i < a.nc) {
    const DevChunk c = a.cols[0].chunks[i];
    a.walk[i] = ChunkWalk{ent_start(c.codec), 0, 0, 0, INT32_MAX, 0};
    a.ev[i] = ~0ull;
  }
}

namespace launch {

void xindex_init(hipStream_t s, const XIndexArgs &a, uint32_t n_spans) {
  const uint32_t n = max(n_spans, a.nc);
  if (n) hipLaunchKernelGGL(k_xinit, dim3((n + 255) / 256), dim3(256), 0, s, a, n_spans);
}
void xindex_cands(hipStream_t s, const XIndexArgs &a) {
  if (a.n_groups) hipLaunchKernelGGL(k_xcands, dim3((a.n_groups + 255) / 256), dim3(256), 0, s, a);
}
void xindex_walk(hipStream_t s, const XIndexArgs &a, uint32_t stop, bool blocks) {
  if (a.nc) hipLaunchKernelGGL(k_xwalk, dim3((a.nc + 3) / 4), dim3(256), 0, s, a, stop, (uint32_t)blocks);
}
void xindex_emit(hipStream_t s, const XIndexArgs &a) {
  if (a.n_groups) hipLaunchKernelGGL(k_xemit, dim3((a.n_groups + 255) / 256), dim3(256), 0, s, a);
  if (a.nc) hipLaunchKernelGGL(k_xresolve, dim3((a.nc + 255) / 256), dim3(256), 0, s, a);
}

}  // namespace launch
}  // namespace ebi
