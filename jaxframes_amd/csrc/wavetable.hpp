// Groupby for small key ranges (max - min < 2048) with wave-private
// accumulator tables in LDS.
//
// Each wave streams a contiguous run of rows and folds every row into its own
// LDS table indexed by digit = ordered(key) - min: no barrier and no global
// atomic in the main loop, one HBM read of keys and values. For ranges of at
// most 256 digits the wave's table is split into LG lane-group tables (lane l
// uses table l % LG), so few distinct keys do not make the lanes of a wave
// collide.
//
// A table entry holds only the compensated sum (and min / max when asked
// for): the row count of every digit is already known exactly from the
// histogram pass, so the tables carry no counts.
//
// Lanes of one wave that hit the same entry in the same instruction are
// serialised by a tag round: every pending lane does ds_max(tag[d], sq|lane)
// then reads tag[d] and the entry back; the lane whose tag survived (the
// highest lane) owns the entry for this round and writes it, the others retry
// in the next round. sq grows every round, so a tag never matches a stale
// round. Tags live in their own packed array so that the atomics of a wave
// spread over all LDS banks. The fold order of every entry is a fixed
// function of the row layout, so results are reproducible run to run.
//
// Each lane folds runs of equal digits among its own consecutive rows in
// registers and writes the table only when the digit changes, so clustered or
// sorted keys cost register adds, not tag rounds.
//
// At the end the block's wave / lane-group tables are merged in a fixed order
// into one partial per (block, digit); k_tile_combine merges the partials in
// block order.
#pragma once

#include <type_traits>

#include "common.hpp"
#include "segreduce.hpp"

namespace jfx {

// Accumulator of one digit: compensated sum, and min / max when MM.
template <typename V, bool MM> struct WAcc;
template <typename V> struct alignas(2 * sizeof(V)) WAcc<V, false> {
  V s, c;
};
template <typename V> struct alignas(4 * sizeof(V) <= 16 ? 4 * sizeof(V) : 16) WAcc<V, true> {
  V s, c, mn, mx;
};

// Lane-group tables per wave for a range of R digits in a TE-entry table:
// each group table needs R entries (a power of two at least R).
__host__ __device__ inline uint32_t wt_lane_groups(uint32_t R, uint32_t te) {
  if (te != 1024u) return 1u;
  return R <= 16u ? 64u : R <= 64u ? 16u : R <= 256u ? 4u : 1u;
}

template <typename K, typename V, bool MM>
struct WaveTabArgs {
  const K* keys;
  const V* vals;
  K sub;                   // raw bits of the minimum key
  uint32_t R;              // digits in use
  uint32_t lg;             // lane-group tables per wave (1, 4, 16 or 64)
  uint32_t rt;             // entries per lane-group table (a power of two >= R)
  uint64_t n;
  uint64_t rows_per_wave;  // multiple of the wave step
  int vec;                 // keys and values are 16-byte aligned
  AggState<V, MM>* part;   // [gridDim.x][R]
  uint32_t* err;           // device error word
  const uint32_t* sel;     // skew selector: run only when it is zero (null: always run)
};

template <typename V, bool MM>
__device__ __forceinline__ void wacc_clear(WAcc<V, MM>& e) {
  e.s = V(0);
  e.c = V(0);
  if constexpr (MM) {
    e.mn = ValTraits<V>::highest();
    e.mx = ValTraits<V>::lowest();
  }
}

// e := e (earlier rows) then r (later rows); the AggState merge rule.
template <typename V, bool MM>
__device__ __forceinline__ void wacc_merge(WAcc<V, MM>& e, const AggState<V, MM>& r) {
  if constexpr (ValTraits<V>::is_float) {
    V err;
    const V s = two_sum(e.s, r.s, err);
    e.c = e.c + r.c + err;
    e.s = s;
  } else {
    using U = typename ValTraits<V>::U;
    e.s = (V)((U)e.s + (U)r.s);
  }
  if constexpr (MM) {
    e.mn = vmin(e.mn, r.mn);
    e.mx = vmax(e.mx, r.mx);
  }
}

template <typename V, bool MM>
__device__ __forceinline__ AggState<V, MM> wacc_state(const WAcc<V, MM>& e) {
  AggState<V, MM> r = AggState<V, MM>::identity();
  r.s = e.s;
  r.c = e.c;
  if constexpr (MM) {
    r.mn = e.mn;
    r.mx = e.mx;
  }
  return r;
}

__device__ __forceinline__ void lds_compiler_fence() { asm volatile("" ::: "memory"); }

typedef uint32_t wt_u32x2 __attribute__((ext_vector_type(2)));

// Reads tag[d] and acc[d] with both reads in flight before one wait (for the
// 8-byte accumulator; the low 32 bits of a generic pointer to LDS are the LDS
// address). Other accumulator sizes use ordinary loads.
template <typename V, bool MM>
__device__ __forceinline__ void wt_read(const uint32_t* tag, const WAcc<V, MM>* acc, uint32_t& t,
                                        WAcc<V, MM>& e) {
  if constexpr (sizeof(WAcc<V, MM>) == 8) {
    wt_u32x2 w;
    asm volatile(
        "ds_read_b32 %0, %2\n\t"
        "ds_read_b64 %1, %3\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(t), "=&v"(w)
        : "v"((uint32_t)(uintptr_t)tag), "v"((uint32_t)(uintptr_t)acc)
        : "memory");
    __builtin_memcpy(&e, &w, 8);
  } else {
    t = *tag;
    e = *acc;
  }
}

// Write the pending states of the lanes with `pend` set into the wave table,
// one tag round per conflict level. The loop runs on the wave-uniform mask of
// pending lanes; `sq` is the round counter pre-shifted by 6.
template <typename V, bool MM>
__device__ __forceinline__ void wt_flush(uint32_t* tags, WAcc<V, MM>* accs, uint32_t slot,
                                         const AggState<V, MM>& st, bool pend, uint32_t& sq,
                                         uint32_t lane) {
  uint64_t pm = ballot64(pend);
  while (pm != 0ull) {
    sq += 64u;
    const uint32_t my = sq | lane;
    // this lane's bit of the uniform mask, as a predicate (no VALU)
    const bool pl = __builtin_amdgcn_inverse_ballot_w64(pm);
    uint32_t t;      // read only where pl (win is false elsewhere)
    WAcc<V, MM> e;
    if (pl) {
      atomicMax(&tags[slot], my);
      lds_compiler_fence();
      wt_read(&tags[slot], &accs[slot], t, e);
    }
    const bool hit = t == my;
    const bool win = pl && hit;
    // ballot of the bare compare (a lane mask with no VALU); lanes outside pm
    // have their bit clear already, whatever their t
    pm &= ~ballot64(hit);
    if (win) {
      wacc_merge(e, st);
      accs[slot] = e;
    }
    lds_compiler_fence();
  }
}

template <typename T>
__device__ __forceinline__ void load4(const T* p, T (&o)[4]) {
  if constexpr (sizeof(T) == 4) {
    const uint4 v = *reinterpret_cast<const uint4*>(p);
    o[0] = __builtin_bit_cast(T, v.x);
    o[1] = __builtin_bit_cast(T, v.y);
    o[2] = __builtin_bit_cast(T, v.z);
    o[3] = __builtin_bit_cast(T, v.w);
  } else {
    const ulonglong2 a = reinterpret_cast<const ulonglong2*>(p)[0];
    const ulonglong2 b = reinterpret_cast<const ulonglong2*>(p)[1];
    o[0] = __builtin_bit_cast(T, a.x);
    o[1] = __builtin_bit_cast(T, a.y);
    o[2] = __builtin_bit_cast(T, b.x);
    o[3] = __builtin_bit_cast(T, b.y);
  }
}

// Rows per lane per step. 16 consecutive rows: runs of one key in the input
// (clustered or sorted data) stay in registers.
constexpr int WT_RPL = 16;

template <typename T>
__device__ __forceinline__ void wt_load_rows(const T* p, T (&o)[WT_RPL]) {
#pragma unroll
  for (int q = 0; q < WT_RPL; q += 4) {
    T t[4];
    load4(p + q, t);
#pragma unroll
    for (int i = 0; i < 4; ++i) o[q + i] = t[i];
  }
}

// No run yet (an impossible digit).
constexpr uint32_t WT_NO_RUN = 0xFFFFFFFFu;

// Fold one row into the lane's run; the run it replaces (if any) is returned
// as a pending state for the table.
template <typename K, typename V, bool MM>
__device__ __forceinline__ void wt_fold(uint32_t dmask, K key, V x, K rawmin, uint32_t& run_d,
                                        AggState<V, MM>& run, AggState<V, MM>& out,
                                        uint32_t& out_d, bool& out_pend) {
  // ordered(k) - ordered(min) == k - min modulo 2^bits for the integer
  // orders; the mask keeps a key outside [min, max] inside the table
  const uint32_t d = (uint32_t)(key - rawmin) & dmask;
  const bool same = d == run_d;
  const uint32_t old_d = run_d;
  out = run;
  out_d = run_d & dmask;
  // run := same ? run + x : {x} (no add at all when no lane extends a run)
  if constexpr (ValTraits<V>::is_float) {
    if (ballot64(same) != 0ull) {
      V err;
      const V t = two_sum(run.s, x, err);
      run.s = same ? t : x;
      run.c = same ? run.c + err : V(0);
    } else {
      run.s = x;
      run.c = V(0);
    }
  } else {
    using U = typename ValTraits<V>::U;
    run.s = same ? (V)((U)run.s + (U)x) : x;
  }
  if constexpr (MM) {
    run.mn = same ? vmin(run.mn, x) : x;
    run.mx = same ? vmax(run.mx, x) : x;
  }
  run_d = d;
  out_pend = d != old_d && old_d != WT_NO_RUN;
}

// Main loop: whole steps of 64 lanes x WT_RPL consecutive rows, the next step
// loaded before the current one is folded. The partial last step of a range
// goes row by row.
template <typename K, typename V, bool MM, int TE, int W, bool VEC>
__global__ void __launch_bounds__(W * 64) k_wave_table(WaveTabArgs<K, V, MM> a) {
  using A = WAcc<V, MM>;
  using St = AggState<V, MM>;
  constexpr uint64_t STEP = 64u * WT_RPL;
  __shared__ A s_acc[W * TE];
  __shared__ uint32_t s_tag[W * TE];
  if (a.sel && *a.sel != 0u) return;

  // wave-uniform (readfirstlane): every range bound below stays scalar
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t lane = lane_id();
  A* accs = s_acc + wid * TE;
  uint32_t* tags = s_tag + wid * TE;
  for (uint32_t i = lane; i < (uint32_t)TE; i += 64) {
    wacc_clear(accs[i]);
    tags[i] = 0u;
  }

  const uint32_t gbase = (lane & (a.lg - 1u)) * a.rt;
  accs += gbase;
  tags += gbase;
  const uint32_t dmask = a.rt - 1u;
  const uint64_t wave = (uint64_t)blockIdx.x * W + wid;
  const uint64_t wbeg = wave * a.rows_per_wave;
  const uint64_t wend = (wbeg + a.rows_per_wave) < a.n ? (wbeg + a.rows_per_wave) : a.n;
  const K rawmin = a.sub;

  uint32_t sq = 0u;
  uint32_t run_d = WT_NO_RUN;
  St run = St::identity();

  uint64_t base = wbeg;
  if (VEC && wbeg + STEP <= wend) {
    // the loads are unconditional (clamped to the last whole step) so that
    // the in-order wait counters let the next step stay in flight
    const uint64_t last = wbeg + ((wend - wbeg) / STEP - 1) * STEP;
    K ka[WT_RPL], kb[WT_RPL];
    V va[WT_RPL], vb[WT_RPL];
    auto load = [&](uint64_t b, K (&kk)[WT_RPL], V (&vv)[WT_RPL]) {
      const uint64_t bc = (b < last ? b : last) + (uint64_t)lane * WT_RPL;
      wt_load_rows(a.keys + bc, kk);
      wt_load_rows(a.vals + bc, vv);
    };
    auto fold = [&](K (&kk)[WT_RPL], V (&vv)[WT_RPL]) {
#pragma unroll
      for (int i = 0; i < WT_RPL; ++i) {
        St ps;
        uint32_t od;
        bool pp;
        wt_fold(dmask, kk[i], vv[i], rawmin, run_d, run, ps, od, pp);
        wt_flush(tags, accs, od, ps, pp, sq, lane);
        __builtin_amdgcn_sched_barrier(0);
      }
    };
    load(base, ka, va);
    __builtin_amdgcn_sched_barrier(0);
    load(base + STEP, kb, vb);
    __builtin_amdgcn_sched_barrier(0);
    while (true) {
      fold(ka, va);
      base += STEP;
      if (base > last) break;
      load(base + STEP, ka, va);
      __builtin_amdgcn_sched_barrier(0);
      fold(kb, vb);
      base += STEP;
      if (base > last) break;
      load(base + STEP, kb, vb);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  // the lane's last run, then rows [base, wend) one row per lane per
  // iteration, each written at once
  wt_flush(tags, accs, run_d & dmask, run, run_d != WT_NO_RUN, sq, lane);
  for (uint64_t r0 = base; r0 < wend; r0 += 64) {
    const uint64_t r = r0 + lane;
    const bool valid = r < wend;
    St one = St::identity();
    uint32_t d = 0u;
    if (valid) {
      one.add(a.vals[r]);
      d = (uint32_t)(a.keys[r] - rawmin) & dmask;
    }
    wt_flush(tags, accs, d, one, valid, sq, lane);
  }

  __syncthreads();
  for (uint32_t d = threadIdx.x; d < a.R; d += W * 64) {
    St acc = St::identity();
    for (int w = 0; w < W; ++w)
      for (uint32_t g = 0; g < a.lg; ++g) acc.merge(wacc_state(s_acc[w * TE + g * a.rt + d]));
    a.part[(size_t)blockIdx.x * a.R + d] = acc;
  }
}

// ---- sums without min / max: tables updated by LDS atomics -------------------
//
// Without min / max an entry is only a sum, and one LDS atomic add per
// flushed run replaces the tag round: a double per digit for float32 values
// (ds_add_f64; measured ~27 cycles per wave instruction on gfx950 against
// ~190 for ds_add_f32 and ~80 for a tag round with its reads and write), the
// value type's unsigned integer for integer values (ds_add_u32 / u64, exact
// wrapping sums). Runs are folded in the same accumulator type in registers.
// Adding float32 values in double rounds each add at 2^-53 of the partial
// sum, no worse than the compensated float pair of the tag-round kernel.
// Lanes of one instruction that add to the same entry are serialised by the
// LDS atomic unit in a fixed order and the tables are wave-private, so the
// result is still a fixed function of the row layout. float64 values keep the
// tag-round kernel (their compensated double-double sums need two words).
template <typename V> struct WaAcc {
  using T = typename ValTraits<V>::U;  // integer values: wrapping unsigned sums
};
template <> struct WaAcc<float> {
  using T = double;
};

template <typename K, typename V, int TE, int W, bool VEC>
__global__ void __launch_bounds__(W * 64) k_wave_atomic(WaveTabArgs<K, V, false> a) {
  static_assert(!std::is_same<V, double>::value, "float64 sums use k_wave_table");
  using T = typename WaAcc<V>::T;
  using St = AggState<V, false>;
  constexpr uint64_t STEP = 64u * WT_RPL;
  __shared__ T s_acc[W * TE];
  if (a.sel && *a.sel != 0u) return;

  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t lane = lane_id();
  T* accs = s_acc + wid * TE;
  for (uint32_t i = lane; i < (uint32_t)TE; i += 64) accs[i] = T(0);
  lds_compiler_fence();

  accs += (lane & (a.lg - 1u)) * a.rt;
  const uint32_t dmask = a.rt - 1u;
  const uint64_t wave = (uint64_t)blockIdx.x * W + wid;
  const uint64_t wbeg = wave * a.rows_per_wave;
  const uint64_t wend = (wbeg + a.rows_per_wave) < a.n ? (wbeg + a.rows_per_wave) : a.n;
  const K rawmin = a.sub;

  uint32_t run_d = WT_NO_RUN;
  T run = T(0);
  auto flush = [&](uint32_t d, T v, bool pend) {
    if (pend) __hip_atomic_fetch_add(&accs[d], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  };
  // fold one row into the lane's run, flushing the run it ends
  auto row = [&](K key, V x) {
    const uint32_t d = (uint32_t)(key - rawmin) & dmask;
    const T xv = (T)x;
    flush(run_d & dmask, run, d != run_d && run_d != WT_NO_RUN);
    run = d == run_d ? run + xv : xv;
    run_d = d;
  };

  uint64_t base = wbeg;
  if (VEC && wbeg + STEP <= wend) {
    // the loads are unconditional (clamped to the last whole step) so that
    // the in-order wait counters let the next step stay in flight
    const uint64_t last = wbeg + ((wend - wbeg) / STEP - 1) * STEP;
    K ka[WT_RPL], kb[WT_RPL];
    V va[WT_RPL], vb[WT_RPL];
    auto load = [&](uint64_t b, K (&kk)[WT_RPL], V (&vv)[WT_RPL]) {
      const uint64_t bc = (b < last ? b : last) + (uint64_t)lane * WT_RPL;
      wt_load_rows(a.keys + bc, kk);
      wt_load_rows(a.vals + bc, vv);
    };
    auto fold = [&](K (&kk)[WT_RPL], V (&vv)[WT_RPL]) {
#pragma unroll
      for (int i = 0; i < WT_RPL; ++i) row(kk[i], vv[i]);
    };
    load(base, ka, va);
    __builtin_amdgcn_sched_barrier(0);
    load(base + STEP, kb, vb);
    __builtin_amdgcn_sched_barrier(0);
    while (true) {
      fold(ka, va);
      base += STEP;
      if (base > last) break;
      load(base + STEP, ka, va);
      __builtin_amdgcn_sched_barrier(0);
      fold(kb, vb);
      base += STEP;
      if (base > last) break;
      load(base + STEP, kb, vb);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  // the lane's last run, then rows [base, wend) one row per lane per
  // iteration, each added at once
  flush(run_d & dmask, run, run_d != WT_NO_RUN);
  for (uint64_t r0 = base; r0 < wend; r0 += 64) {
    const uint64_t r = r0 + lane;
    const bool valid = r < wend;
    uint32_t d = 0u;
    T v = T(0);
    if (valid) {
      v = (T)a.vals[r];
      d = (uint32_t)(a.keys[r] - rawmin) & dmask;
    }
    flush(d, v, valid);
  }

  __syncthreads();
  for (uint32_t d = threadIdx.x; d < a.R; d += W * 64) {
    T acc = T(0);
    for (int w = 0; w < W; ++w)
      for (uint32_t g = 0; g < a.lg; ++g) acc += s_acc[w * TE + g * a.rt + d];
    St st = St::identity();
    if constexpr (std::is_same<V, float>::value) {
      // the double as a float pair s + c (c the rounding error of s, 0 when
      // s is not finite)
      st.s = (float)acc;
      st.c = __builtin_isfinite(st.s) ? (float)(acc - (double)st.s) : 0.0f;
    } else {
      st.s = (V)acc;
    }
    a.part[(size_t)blockIdx.x * a.R + d] = st;
  }
}

}  // namespace jfx
