// One-pass groupby for integer keys whose range turns out to be below the
// table width TE (1024, or 2048 for a retry): sums, counts and means (no min
// / max) of float32 or integer values.
//
// Distinct integers inside any window of TE consecutive values have distinct
// residues mod TE, so wave tables indexed by key mod TE need no minimum key in
// advance. k_onepass_table reads keys and values once: each lane folds runs
// of equal residues in registers (as the wave-table kernel does) and adds
// every finished run with two LDS atomics: the sum into its wave's table
// (ds_add_f64 / integer add; wave-private so that the float add order is
// fixed) and the row count into its block's table (ds_add_u32: integer adds
// commute, so the 4 waves share one count table, 36 KB of LDS per block and 4
// blocks per CU at TE = 1024). It also tracks the ordered min / max of every
// key it reads. When 64 keys sampled over the whole input span few values,
// the table is split into lane-group tables of rt < TE residues (op_residues).
// k_onepass_merge and k_onepass_final merge the block tables per residue in
// block order, take the global min / max, and when max - min < rt write the
// groups in key order (digit d = ordered(key) - min, residue = (min + d) mod
// rt). Otherwise they report a miss with a lower bound of the key span, and
// the caller retries once with wider tables or runs the statistics-first
// path. The result of a hit is a fixed function of the row layout.
//
// Early exit: a wave first reads the 64 sampled keys, and checks its own keys
// again after its first step of rows; when either spans rt or more values the
// final check cannot pass, so the wave records that span in its block's min /
// max and skips its rows, and the call's stop word (epoch-tagged: no clearing
// needed) is raised once, for the other waves, which poll it once per step,
// and for the merge kernels.
#pragma once

#include "blockscan.hpp"
#include "common.hpp"
#include "runtime.hpp"
#include "segreduce.hpp"
#include "tilereduce.hpp"
#include "wavetable.hpp"

namespace jfx {

constexpr uint32_t OP_TE = 1024;  // table entries (residues) per wave; 2048 for the wider retry
constexpr int OP_RPL = 16;        // rows per lane per step of k_onepass_table
constexpr int OP_W = 4;           // waves per block
constexpr uint32_t OP_NO_RUN = 0xFFFFFFFFu;

// Accumulator of a residue's sum: double for float32 and float64 values
// (float64 with a compensation term, see k_onepass_table), the value type's
// unsigned integer for integer values (wrapping sums).
template <typename V> struct OpT {
  using T = typename WaAcc<V>::T;
};
template <> struct OpT<double> {
  using T = double;
};
template <typename V>
constexpr bool op_dd = std::is_same<V, double>::value;

// Order-preserving unsigned codes of a value for the min / max tables (LDS
// integer atomics; float LDS atomics serialise their lanes). Float NaN codes
// below every value in the min table and above every value in the max table,
// so a NaN anywhere in a group makes its min and max NaN (numpy's minimum /
// maximum reductions); -0.0 and +0.0 share one code (decoded as +0.0).
template <typename V> struct MmCode {
  using U = typename ValTraits<V>::U;
  static __device__ __forceinline__ U lo(V x) {
    if constexpr (std::is_same<V, float>::value) {
      const uint32_t b = __float_as_uint(x);
      return (b & 0x7FFFFFFFu) > 0x7F800000u ? 0u : ordered(b, ORD_FLOAT);
    } else if constexpr (std::is_same<V, double>::value) {
      const uint64_t b = (uint64_t)__double_as_longlong(x);
      return (b & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull ? 0ull : ordered(b, ORD_FLOAT);
    } else if constexpr (std::is_signed<V>::value) {
      return (U)x ^ ((U)1 << (sizeof(U) * 8 - 1));
    } else {
      return (U)x;
    }
  }
  static __device__ __forceinline__ U hi(V x) {
    if constexpr (std::is_same<V, float>::value) {
      const uint32_t b = __float_as_uint(x);
      return (b & 0x7FFFFFFFu) > 0x7F800000u ? 0xFFFFFFFFu : ordered(b, ORD_FLOAT);
    } else if constexpr (std::is_same<V, double>::value) {
      const uint64_t b = (uint64_t)__double_as_longlong(x);
      return (b & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull ? ~0ull : ordered(b, ORD_FLOAT);
    } else {
      return lo(x);
    }
  }
  static __device__ __forceinline__ V decode(U u) {
    if constexpr (std::is_same<V, float>::value) {
      if (u == 0u || u == 0xFFFFFFFFu) return __uint_as_float(0x7FC00000u);
      return __uint_as_float((u & 0x80000000u) ? (u ^ 0x80000000u) : ~u);
    } else if constexpr (std::is_same<V, double>::value) {
      if (u == 0ull || u == ~0ull) return __longlong_as_double(0x7FF8000000000000ll);
      const uint64_t sgn = 0x8000000000000000ull;
      return __longlong_as_double((long long)((u & sgn) ? (u ^ sgn) : ~u));
    } else if constexpr (std::is_signed<V>::value) {
      return (V)(u ^ ((U)1 << (sizeof(U) * 8 - 1)));
    } else {
      return (V)u;
    }
  }
};

// One residue of a block table: the sum, the row count and (MM) the min /
// max codes. float32 values keep their double sum as a float pair s + c (c
// the float32 rounding of the remainder, 0 when s is not finite: the pair
// keeps about 48 of the double's 53 significant bits, a relative loss of at
// most 2^-48 of the partial sum, far inside the float32 result's tolerance);
// integer values a wrapping unsigned sum.
// (Plain structs: with a base struct the merge loop's 16 loads were no
// longer kept in flight together.)
template <typename V, bool MM> struct OpPart {
  typename ValTraits<V>::U s;
  uint32_t n;
};
template <typename V> struct OpPart<V, true> {
  typename ValTraits<V>::U s;
  uint32_t n;
  typename ValTraits<V>::U mn, mx;
};
template <> struct OpPart<float, false> {
  float s, c;
  uint32_t n;
};
template <> struct OpPart<float, true> {
  float s, c;
  uint32_t n;
  uint32_t mn, mx;
};
// float64: the compensated pair s + c itself
template <> struct OpPart<double, false> {
  double s, c;
  uint32_t n;
};
template <> struct OpPart<double, true> {
  double s, c;
  uint32_t n;
  uint64_t mn, mx;
};

// Running state of one residue: the sum in the accumulator type, the count
// and (MM) the min / max codes.
template <typename V, bool MM> struct OpAcc {
  using T = typename OpT<V>::T;
  using U = typename ValTraits<V>::U;
  T s;
  T c;  // float64 values: compensation of s (TwoSum errors)
  uint32_t n;
  U mn, mx;
  __device__ __forceinline__ void init() {
    s = T(0);
    c = T(0);
    n = 0u;
    mn = ~U(0);
    mx = U(0);
  }
  // s, c := s, c then (ps, pc)
  __device__ __forceinline__ void add_pair(T ps, T pc) {
    T e;
    const T t = two_sum(s, ps, e);
    c = c + pc + e;
    s = t;
  }
  __device__ __forceinline__ void add(const OpPart<V, MM>& p) {
    if constexpr (std::is_same<V, float>::value) s += (double)p.s + (double)p.c;
    else if constexpr (op_dd<V>) add_pair(p.s, p.c);
    else s += p.s;
    n += p.n;
    if constexpr (MM) {
      mn = p.mn < mn ? p.mn : mn;
      mx = p.mx > mx ? p.mx : mx;
    }
  }
  __device__ __forceinline__ OpPart<V, MM> pack() const {
    OpPart<V, MM> p;
    if constexpr (std::is_same<V, float>::value) {
      p.s = (float)s;
      p.c = __builtin_isfinite(p.s) ? (float)(s - (double)p.s) : 0.0f;
    } else if constexpr (op_dd<V>) {
      p.s = s;
      p.c = c;
    } else {
      p.s = s;
    }
    p.n = n;
    if constexpr (MM) {
      p.mn = mn;
      p.mx = mx;
    }
    return p;
  }
};

template <typename K, typename V, bool MM>
struct OnePassArgs {
  const K* keys;             // raw key bits (integer keys)
  const V* vals;
  uint64_t n;
  int oc;                    // key order code
  uint32_t epoch;            // this call's tag for the stop word
  uint32_t rt_min;           // residues per lane-group table at least (TE: no lane groups)
  OpPart<V, MM>* part;       // [gridDim.x][TE] block tables
  uint64_t* mm;              // [gridDim.x][2] ordered min and max of the block's keys
  uint32_t* ctl;             // [0] stop word
};

template <typename K>
__device__ __forceinline__ K wave_min_k(K v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const K o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}
template <typename K>
__device__ __forceinline__ K wave_max_k(K v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const K o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

// Residues per lane-group table for a 64-key sample spanning `span` values
// (2x margin): few distinct keys would make the lanes of a wave collide on
// one entry (same-address LDS atomics serialise), so the TE entries are
// split into TE / rt tables and lane l uses table l % (TE / rt). The
// final check then requires max - min < rt. Both kernels derive rt from the
// same 64 keys, so every wave agrees without communicating.
template <uint32_t TE, typename K>
__device__ __forceinline__ uint32_t op_residues(K span) {
  return span < (K)8 ? 16u : span < (K)32 ? 64u : span < (K)128 ? 256u : TE;
}

// Span (max - min of the ordered keys) of the 64 keys at rows lane * n / 64.
template <typename K>
__device__ __forceinline__ K op_sample_span(const K* keys, uint64_t n, int oc, uint32_t lane) {
  const K o = ordered(keys[(uint64_t)lane * n / 64u], oc);
  return wave_max_k(o) - wave_min_k(o);
}

// Rows per lane per load of the table kernel's steps: E rows of the wider of
// the key and value types fill 16 bytes.
template <typename K, typename V>
constexpr int OP_E = 16 / (sizeof(K) > sizeof(V) ? sizeof(K) : sizeof(V));
typedef uint32_t op_v4u __attribute__((ext_vector_type(4)));
typedef uint32_t op_v2u __attribute__((ext_vector_type(2)));
// One step's 16 rows of a lane: rows [c*64*E + lane*E, +E) of the step for c
// = 0 .. 16/E - 1 into o[c*E .. c*E+E). Each load instruction covers 64
// consecutive pieces, so it consumes whole cache lines, and the loads are
// non-temporal (every row is read once). Measured at 100M int32 keys and
// float32 values: 138-143 us, against 151-154 us for 16 consecutive rows per
// lane with ordinary loads and 234 us with non-temporal ones (each such
// instruction touches 64 lines in part). Runs of equal keys then fold in
// registers over E consecutive rows instead of 16: keys in runs of 64 rows
// measured 186 us against 152 us (16 lanes flush one residue together).
template <int E, typename T, int N>
__device__ __forceinline__ void op_load_coal(const T* p, uint32_t lane, T (&o)[N]) {
  constexpr int B = E * sizeof(T);
#pragma unroll
  for (int c = 0; c < N / E; ++c) {
    const T* src = p + (uint64_t)c * 64 * E + (uint64_t)lane * E;
    T t[E];
    if constexpr (B == 16) {
      const op_v4u* q = reinterpret_cast<const op_v4u*>(src);
      const op_v4u v = __builtin_nontemporal_load(q);
      __builtin_memcpy(t, &v, 16);
    } else {
      const op_v2u* q = reinterpret_cast<const op_v2u*>(src);
      const op_v2u v = __builtin_nontemporal_load(q);
      __builtin_memcpy(t, &v, 8);
    }
#pragma unroll
    for (int i = 0; i < E; ++i) o[c * E + i] = t[i];
  }
}

// W waves per block (the workgroup's wave tables are merged into one block
// table at its end). NC (float32 sums without counts, min / max or means):
// no count table. The
// wave tables start at -0.0 and every run is added as run + 0.0, which maps
// -0.0 to +0.0 and keeps every other value: a round-to-nearest sum of
// addends that are not all -0.0 is never -0.0, so an entry still holding
// -0.0 saw no row, and the block table's count field carries that presence
// (0 or 1). The sums are those of tables started at +0.0 (adding -0.0 to a
// sum changes nothing). This saves the count atomic of every run, a third
// of the LDS atomic time, and 4 KB of LDS per block.
template <typename K, typename V, bool MM, bool VEC, uint32_t TE, bool NC = false, int W = OP_W>
__global__ void __launch_bounds__(W * 64) k_onepass_table(OnePassArgs<K, V, MM> a) {
  using T = typename OpT<V>::T;
  using U = typename ValTraits<V>::U;
  using C = MmCode<V>;
  constexpr bool DD = op_dd<V>;
  static_assert(!NC || (std::is_same<V, float>::value && !MM), "count-free tables: float32 sums only");
  constexpr uint64_t STEP = 64u * OP_RPL;
  __shared__ T s_acc[W * TE];
  // float64 values: wave-private compensation of each sum entry
  __shared__ T s_comp[DD ? W * TE : 1];
  __shared__ uint32_t s_cnt[NC ? 1 : TE];
  // min / max codes: order-independent like the counts, one table per block
  __shared__ U s_min[MM ? TE : 1], s_max[MM ? TE : 1];
  __shared__ K s_mn[W], s_mx[W];
  __shared__ uint32_t s_stop;
  if (threadIdx.x == 0) s_stop = 0u;
  for (uint32_t i = threadIdx.x; i < TE; i += W * 64) {
    if constexpr (!NC) s_cnt[i] = 0u;
    if constexpr (MM) {
      s_min[i] = ~U(0);
      s_max[i] = U(0);
    }
  }

  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t lane = lane_id();
  for (uint32_t i = lane; i < TE; i += 64) {
    s_acc[wid * TE + i] = NC ? T(-0.0) : T(0);
    if constexpr (DD) s_comp[wid * TE + i] = T(0);
  }
  __syncthreads();

  K omin = ~K(0), omax = K(0);
  auto see = [&](K key) {
    const K o = ordered(key, a.oc);
    omin = o < omin ? o : omin;
    omax = o > omax ? o : omax;
  };
  // 64 keys spread over the whole input (index lane * n / 64 < n) choose the
  // lane-group tables (every wave reads the same keys)
  const K sspan = op_sample_span(a.keys, a.n, a.oc, lane);
  const uint32_t rt0 = op_residues<TE>(sspan);
  const uint32_t rt = rt0 > a.rt_min ? rt0 : a.rt_min;
  const uint32_t M = rt - 1u;
  // every wave chose the same rt; the final merge reads it from ctl[1]
  if (blockIdx.x == 0 && threadIdx.x == 0) a.ctl[1] = rt;
  // lane group g holds residue r at entry (r ^ g) mod rt of its table: with
  // tables rt entries apart, an unswizzled residue would put the lanes of a
  // wave on a few LDS banks
  const uint32_t gsw = lane & (TE / rt - 1u);
  T* accs = s_acc + wid * TE + gsw * rt;
  T* comps = s_comp + (DD ? wid * TE + gsw * rt : 0u);
  uint32_t* cnts = s_cnt + gsw * rt;
  U* mins = s_min + (MM ? gsw * rt : 0u);
  U* maxs = s_max + (MM ? gsw * rt : 0u);
  // the wave's keys so far span rt or more values: the final check fails
  auto too_wide = [&]() -> bool {
    const K lo = wave_min_k(omin), hi = wave_max_k(omax);
    return hi >= lo && hi - lo >= (K)rt;
  };
  auto raise_stop = [&]() {
    if (lane == 0) __hip_atomic_store(&a.ctl[0], a.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  };
  auto stopped = [&]() -> bool {
    return __hip_atomic_load(&a.ctl[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == a.epoch;
  };

  // The sample already spanning TE or more values is a miss. Every wave
  // reaches the same verdict, so one wave raises the stop word: same-address
  // stores from thousands of waves serialise at the coherence point (~30 ns
  // each, measured 130 us for a miss).
  see(a.keys[(uint64_t)lane * a.n / 64u]);
  bool stop = sspan >= (K)TE;
  if (stop && blockIdx.x == 0 && wid == 0) raise_stop();

  // Steps of 1024 rows are dealt out grid-stride (wave w takes steps w, w +
  // W, w + 2W, ... of the W waves), so that at any time the waves read one
  // window of consecutive steps: 137-139 us at 100M rows against 140-141 us
  // for one contiguous range of steps per wave.
  const uint64_t wave = (uint64_t)blockIdx.x * W + wid;
  const uint64_t nwaves = (uint64_t)gridDim.x * W;
  const uint64_t nsteps = a.n / STEP;
  const uint64_t SW = STEP * nwaves;

  uint32_t run_r = OP_NO_RUN;
  T run = T(0);
  T run_c = T(0);  // float64: the run's compensation
  uint32_t run_n = 0u;
  U run_lo = ~U(0), run_hi = U(0);
  // float64 values add the run with ds_add_rtn_f64, which returns the sum it
  // was added to, so TwoSum(old, v) is the exact rounding error of that add;
  // it goes, with the run's own compensation vc, into the compensation entry
  auto flush = [&](uint32_t r, T v, T vc, uint32_t c, U lo, U hi, bool pend) {
    if (pend) {
      if constexpr (DD) {
        const T old = __hip_atomic_fetch_add(&accs[r], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        T e;
        (void)two_sum(old, v, e);
        __hip_atomic_fetch_add(&comps[r], vc + e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      } else if constexpr (NC) {
        (void)vc;
        (void)c;
        __hip_atomic_fetch_add(&accs[r], v + T(0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      } else {
        (void)vc;
        __hip_atomic_fetch_add(&accs[r], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      }
      if constexpr (!NC)
        __hip_atomic_fetch_add(&cnts[r], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      if constexpr (MM) {
        __hip_atomic_fetch_min(&mins[r], lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_max(&maxs[r], hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
  };
  auto row = [&](K key, V x) {
    see(key);
    const uint32_t r = ((uint32_t)key ^ gsw) & M;
    flush(run_r, run, run_c, run_n, run_lo, run_hi, r != run_r && run_r != OP_NO_RUN);
    const bool same = r == run_r;
    if constexpr (DD) {
      T e;
      const T t = two_sum(run, (T)x, e);
      run = same ? t : (T)x;
      run_c = same ? run_c + e : T(0);
    } else {
      run = same ? run + (T)x : (T)x;
    }
    run_n = same ? run_n + 1u : 1u;
    if constexpr (MM) {
      const U lo = C::lo(x), hi = C::hi(x);
      run_lo = same && run_lo < lo ? run_lo : lo;
      run_hi = same && run_hi > hi ? run_hi : hi;
    }
    run_r = r;
  };

  uint64_t base = wave * STEP;
  if (!stop && VEC && wave < nsteps) {
    // the loads are unconditional (clamped to the last whole step)
    const uint64_t last = base + ((nsteps - 1 - wave) / nwaves) * SW;
    K ka[OP_RPL];
    V va[OP_RPL];
    auto load = [&](uint64_t b, K (&kk)[OP_RPL], V (&vv)[OP_RPL]) {
      const uint64_t bc = (b < last ? b : last);
      op_load_coal<OP_E<K, V>>(a.keys + bc, lane, kk);
      op_load_coal<OP_E<K, V>>(a.vals + bc, lane, vv);
    };
    auto fold = [&](K (&kk)[OP_RPL], V (&vv)[OP_RPL]) {
#pragma unroll
      for (int i = 0; i < OP_RPL; ++i) row(kk[i], vv[i]);
    };
    bool first = true;
    // after a step: the first step's own span, then the stop word
    auto check = [&]() -> bool {
      if (first) {
        first = false;
        if (too_wide()) {
          if (!stopped()) raise_stop();
          return true;
        }
      }
      return stopped();
    };
    // One step of loads in flight per wave (8 KB of 4-byte keys and values):
    // at 100M rows the 4096 waves then keep 32 MB in flight, and the table
    // kernel took 128.1-128.9 us, against 134.1-134.7 us with the next step
    // loaded while folding the current one (16 KB per wave), 132.5 us for
    // single 512-row steps and 133.5 us for single 2048-row steps. The
    // hand-written read stream of the bandwidth probe likewise runs fastest
    // with one 16-byte load in flight per thread.
    load(base, ka, va);
    __builtin_amdgcn_sched_barrier(0);
    while (true) {
      fold(ka, va);
      base += SW;
      if (base > last || (stop = check())) break;
      load(base, ka, va);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  if (!stop) {
    flush(run_r, run, run_c, run_n, run_lo, run_hi, run_r != OP_NO_RUN);
  }
  // rows one per lane: without vector loads every step of the wave, and the
  // rows after the last whole step for the wave that would take the next
  const uint64_t tail_wave = nsteps % nwaves;
  const uint64_t nseg = stop ? 0u : (VEC ? 0u : (wave < nsteps ? (nsteps - 1 - wave) / nwaves + 1 : 0u)) +
                                        (wave == tail_wave && nsteps * STEP < a.n ? 1u : 0u);
  for (uint64_t sg = 0; sg < nseg; ++sg) {
    const bool tail = !VEC ? (wave + sg * nwaves >= nsteps) : true;
    const uint64_t wbeg = tail ? nsteps * STEP : (wave + sg * nwaves) * STEP;
    const uint64_t wend = tail ? a.n : wbeg + STEP;
    for (uint64_t r0 = wbeg; r0 < wend; r0 += 64) {
      const uint64_t r = r0 + lane;
      const bool valid = r < wend;
      K key = K(0);
      V x = V(0);
      if (valid) {
        key = a.keys[r];
        x = a.vals[r];
        see(key);
      }
      flush(((uint32_t)key ^ gsw) & M, (T)x, T(0), 1u, C::lo(x), C::hi(x), valid);
    }
  }

  const K lo = wave_min_k(omin), hi = wave_max_k(omax);
  if (lane == 0) {
    s_mn[wid] = lo;
    s_mx[wid] = hi;
    if (stop) s_stop = 1u;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    K bl = s_mn[0], bh = s_mx[0];
    for (int w = 1; w < W; ++w) {
      bl = s_mn[w] < bl ? s_mn[w] : bl;
      bh = s_mx[w] > bh ? s_mx[w] : bh;
    }
    a.mm[(size_t)blockIdx.x * 2 + 0] = (uint64_t)bl;
    a.mm[(size_t)blockIdx.x * 2 + 1] = (uint64_t)bh;
  }
  if (s_stop) return;  // block-uniform: the result will be a miss
  // residue r < rt: its lane-group tables in (wave, group) order; the
  // entries from rt on are unused
  const uint32_t lg = TE / rt;
  for (uint32_t r = threadIdx.x; r < TE; r += W * 64) {
    OpAcc<V, MM> st;
    st.init();
    if (r < rt) {
      bool seen = false;
      for (int w = 0; w < W; ++w)
        for (uint32_t g = 0; g < lg; ++g) {
          const uint32_t e = w * TE + g * rt + ((r ^ g) & (rt - 1u));
          if constexpr (DD) st.add_pair(s_acc[e], s_comp[e]);
          else st.s += s_acc[e];
          if constexpr (NC) seen |= __double_as_longlong(s_acc[e]) != (long long)0x8000000000000000ull;
        }
      if constexpr (NC) st.n = seen ? 1u : 0u;
      for (uint32_t g = 0; g < (NC ? 0u : lg); ++g) {
        const uint32_t e = g * rt + ((r ^ g) & (rt - 1u));
        st.n += s_cnt[e];
        if constexpr (MM) {
          st.mn = s_min[e] < st.mn ? s_min[e] : st.mn;
          st.mx = s_max[e] > st.mx ? s_max[e] : st.mx;
        }
      }
    }
    a.part[(size_t)blockIdx.x * TE + r] = st.pack();
  }
}

// Result that k_onepass_final publishes to the host as one word (rb_word): a
// status and the groups (on a miss: a lower bound of the key span max - min,
// saturated to 2^30 - 1, which the retry decisions below 2048 do not see).
constexpr uint32_t OP_MISS = 0u, OP_HIT = 1u, OP_CAPACITY = 2u;

__device__ __forceinline__ void op_publish(const ReadBackSlot& rb, uint32_t status, uint32_t value) {
  __hip_atomic_store(rb.seq_slot, rb_word(rb.seq, status, value), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_SYSTEM);
}

// Two-level merge, thread = residue (coalesced reads): block (g, c) of
// k_onepass_merge merges residues [OP_MT * c, OP_MT * (c + 1)) of block
// tables [group * g, group * (g + 1)) in block order into l1[g] (group =
// OP_GROUP, or 16 for the 256 blocks of 16 waves of the count-free tables);
// k_onepass_final (one block) merges the at most OP_L1 level-1 tables in
// order, checks the key range and writes the groups in key order. Each merge
// issues all its loads before the first add: the level-1 blocks are small
// enough to hold 64 states per thread (32 of the 32- and 40-byte states that
// carry 64-bit min / max codes), the final block of 1024 threads holds its 16. (A last-block ticket
// in one kernel measured slower: its release fences write back L2.)
constexpr uint32_t OP_GROUP = 64;
constexpr uint32_t OP_L1 = 16;  // up to 1024 block tables
constexpr uint32_t OP_MT = 128;  // level-1 merge threads per block

// acc, n := acc, n plus p[0], p[stride], ... (count entries, added in order
// in the accumulator type), G loads in flight at a time.
template <uint32_t G, typename V, bool MM>
__device__ __forceinline__ void merge_seq(OpAcc<V, MM>& st, const OpPart<V, MM>* p, size_t stride,
                                          uint32_t count) {
  OpAcc<V, MM> x = st;  // a local copy: accumulating through the reference kept
                        // the compiler from leaving the 16 loads in flight
  for (uint32_t b = 0; b < count; b += G, p += G * stride) {
    const uint32_t m = (count - b) < G ? (count - b) : G;
    OpPart<V, MM> v[G];
#pragma unroll
    for (uint32_t i = 0; i < G; ++i)
      if (i < m) v[i] = p[i * stride];
#pragma unroll
    for (uint32_t i = 0; i < G; ++i)
      if (i < m) x.add(v[i]);
  }
  st = x;
}

// grid (groups, TE / OP_MT); group (at most OP_GROUP) block tables per
// level-1 table
template <typename V, bool MM, uint32_t TE>
__global__ void __launch_bounds__(OP_MT)
k_onepass_merge(const OpPart<V, MM>* part, uint32_t nblocks, uint32_t group, const uint32_t* ctl,
                uint32_t epoch, OpPart<V, MM>* l1) {
  if (ctl[0] == epoch) return;  // stopped: the result is a miss
  constexpr uint32_t G = sizeof(OpPart<V, MM>) <= 24 ? OP_GROUP : OP_GROUP / 2;
  const uint32_t c0 = blockIdx.x * group;
  const uint32_t c1 = (c0 + group) < nblocks ? (c0 + group) : nblocks;
  const uint32_t r = blockIdx.y * OP_MT + threadIdx.x;
  OpAcc<V, MM> st;
  st.init();
  merge_seq<G>(st, part + (size_t)c0 * TE + r, TE, c1 - c0);
  l1[(size_t)blockIdx.x * TE + r] = st.pack();
}

template <typename K, typename V, typename MT, bool MM, uint32_t TE>
__global__ void __launch_bounds__(1024)
k_onepass_final(const OpPart<V, MM>* l1, uint32_t nl1, const uint64_t* mm, uint32_t nblocks,
                const uint32_t* ctl, uint32_t epoch, int oc, K* out_keys, GroupOut<V, MT> out,
                uint64_t capacity, ReadBackSlot rb) {
  using T = typename OpT<V>::T;
  using U = typename ValTraits<V>::U;
  __shared__ K s_lo[16], s_hi[16];
  constexpr bool DD = op_dd<V>;
  __shared__ T s_sum[TE];
  __shared__ T s_c[DD ? TE : 1];
  __shared__ uint32_t s_n[TE];
  __shared__ U s_min[MM ? TE : 1], s_max[MM ? TE : 1];
  __shared__ uint32_t s_gid[TE];
  __shared__ uint32_t s_tmp[1024 / 64 + 1];
  const int tid = threadIdx.x;
  const int wid = tid >> 6;
  const uint32_t lane = lane_id();
  const bool stopped = ctl[0] == epoch;
  // the table kernel's residues per table (loaded with the merge's loads, used
  // after them)
  const uint32_t rt = ctl[1];
  // block min / max loads first, so that they are in flight with the merge's
  K lo = ~K(0), hi = K(0);
  for (uint32_t b = tid; b < nblocks; b += 1024) {
    const K l = (K)mm[(size_t)b * 2], h = (K)mm[(size_t)b * 2 + 1];
    lo = l < lo ? l : lo;
    hi = h > hi ? h : hi;
  }
  // merged without waiting for the stop word (after a stop l1 holds stale
  // states, which a miss never reads out)
  for (uint32_t r = tid; r < TE; r += 1024) {
    OpAcc<V, MM> st;
    st.init();
    merge_seq<OP_L1>(st, l1 + r, TE, nl1);
    s_sum[r] = st.s;
    if constexpr (DD) s_c[r] = st.c;
    s_n[r] = st.n;
    if constexpr (MM) {
      s_min[r] = st.mn;
      s_max[r] = st.mx;
    }
  }

  lo = wave_min_k(lo);
  hi = wave_max_k(hi);
  if (lane == 0) {
    s_lo[wid] = lo;
    s_hi[wid] = hi;
  }
  __syncthreads();
  lo = s_lo[0];
  hi = s_hi[0];
  for (int w = 1; w < 16; ++w) {
    lo = s_lo[w] < lo ? s_lo[w] : lo;
    hi = s_hi[w] > hi ? s_hi[w] : hi;
  }
  const bool hit = !stopped && hi >= lo && hi - lo < (K)rt;
  if (!hit) {
    if (tid == 0) {
      // a lower bound of max - min over all keys (the keys that were read),
      // saturated to 30 bits
      const K span = hi >= lo ? hi - lo : K(0);
      op_publish(rb, OP_MISS, span > (K)0x3FFFFFFFu ? 0x3FFFFFFFu : (uint32_t)span);
    }
    return;
  }
  // digit d = ordered key - lo, in key order; its residue holds its group
  // (TE / 1024 digits per thread)
  const uint32_t span = (uint32_t)(hi - lo);
  auto count_of = [&](uint32_t d) -> uint32_t {
    return d <= span ? s_n[(uint32_t)(lo + (K)d) & (rt - 1u)] : 0u;
  };
  for (uint32_t d = tid; d < TE; d += 1024) s_gid[d] = count_of(d) != 0u ? 1u : 0u;
  __syncthreads();
  const uint32_t G = block_excl_scan_lds<TE, 1024>(s_gid, s_tmp);
  if ((uint64_t)G > capacity) {
    if (tid == 0) op_publish(rb, OP_CAPACITY, G);
    return;
  }
  for (uint32_t d = tid; d < TE; d += 1024) {
    const uint32_t cnt = count_of(d);
    if (cnt == 0u) continue;
    const K o = lo + (K)d;
    const uint32_t r = (uint32_t)o & (rt - 1u);
    T sum = s_sum[r];
    if constexpr (DD) sum = sum + s_c[r];
    const uint32_t g = s_gid[d];
    K raw = o;
    if (oc == ORD_SIGNED) raw ^= (K)1 << (sizeof(K) * 8 - 1);
    if (out_keys) out_keys[g] = raw;
    if (out.sum) out.sum[g] = (V)sum;
    if (out.count) out.count[g] = (int64_t)cnt;
    if constexpr (MM) {
      if (out.mn) out.mn[g] = MmCode<V>::decode(s_min[r]);
      if (out.mx) out.mx[g] = MmCode<V>::decode(s_max[r]);
    }
    if (out.mean) {
      if constexpr (ValTraits<V>::is_float) out.mean[g] = (MT)(sum / (double)cnt);
      else out.mean[g] = (MT)((double)(V)sum / (double)cnt);
    }
  }
  // every thread's outputs are written before the hit is published
  __syncthreads();
  if (tid == 0) op_publish(rb, OP_HIT, G);
}

}  // namespace jfx
