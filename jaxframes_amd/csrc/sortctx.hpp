// Internal sort pipeline shared by the sort, groupby and join entry points.
#pragma once

#include <stdint.h>

#include "runtime.hpp"

namespace jfx {

constexpr int SORT_MAXP = 8;    // most layout passes (64-bit keys, 8-bit digits)
constexpr int SORT_MAXR = 2048; // largest radix
constexpr uint32_t SORT_MAXCHUNKS = 512;

// State of one sort between the statistics pass and the scatter passes.
struct SortCtx {
  int kt = 0;            // key type code
  int kb = 4;            // key bytes
  int oc = 0;            // order code
  int pb = 0;            // payload bytes (0, 4, 8)
  int digit_bits = 8;    // layout digit width for multi-pass sorts
  int shape = 0;         // onesweep tile shape (SORT_SHAPES)
  uint32_t desc_tile = 8192;  // keys per tile the lookback descriptors are carved for
  uint64_t n = 0;
  hipStream_t s = nullptr;
  // chunk geometry of the chunked passes
  uint64_t chunk = 0;
  uint32_t nchunks = 0;
  // statistics of the ordered keys (read back from the device)
  uint64_t mn = 0, mx = 0, band = 0, bor = 0;
  // workspace pieces
  void* alt_keys = nullptr;
  void* alt_pay = nullptr;
  uint32_t* counts = nullptr;   // [nchunks][2048] low-11-bit chunk histograms
  uint32_t* colex = nullptr;    // [nchunks][R] chunk column prefixes
  uint32_t* total = nullptr;    // [2048] digit totals
  uint32_t* hist = nullptr;     // [MAXP][MAXR] layout histograms
  uint32_t* base = nullptr;     // [MAXP][MAXR] layout digit offsets
  unsigned long long* stats = nullptr;  // min, max, and, or
  unsigned long long* stats_part = nullptr;  // [nchunks][4] per-chunk statistics
  uint32_t* maxbin = nullptr;   // [nchunks] largest low-11-bit bin of each chunk
  // sum of the chunks' largest bins (an upper bound of the largest digit
  // count of a small key range); unknown (all ones) for known-range sorts
  uint64_t maxbin_sum = ~0ull;
  // the statistics pass also left every 8-bit layout histogram in hist
  bool hist_done = false;
  uint32_t* counters = nullptr;
  uint32_t* hdr = nullptr;
  uint32_t* desc[2] = {nullptr, nullptr};
  uint32_t* err = nullptr;

  bool single_pass() const { return (mx - mn) < 2048ull; }
  uint32_t range_radix() const { return (uint32_t)(mx - mn) + 1u; }
};

// Onesweep tile shapes (waves x keys per lane): 0 = 32-bit keys 16 x 8,
// 64-bit keys 8 x 16 (8192-key tiles, the default; 64-bit keys with at most
// a 4-byte payload and 8-bit digits take 16 x 16); 1 = 8 x 16 / 16 x 8
// (8192); 2 = 8 x 8 / 4 x 16 (4096); 3 = 4 x 16 / 8 x 8 (4096); 4 = 16 x 16
// for both widths (16384-key tiles, 8-bit digits only: 11-bit digits take
// shape 0). Shapes 1-4 are taken by sorts that carry a payload.
constexpr int SORT_SHAPES = 5;

// small_tiles: carve lookback descriptors for 4096-key tiles (any shape).
uint64_t sort_ws_bytes(int kt, int pay_bytes, uint64_t n, bool small_tiles = false);
int default_digit_bits(int kt, uint64_t n);

// Carve the workspace and run the statistics pass; synchronises the stream
// once to read min / max / AND / OR of the ordered keys into ctx. With
// with_hist (8-bit digits only) the same read also makes the layout
// histograms, for callers that will most likely run a multi-pass sort.
int sort_prepare(SortCtx& ctx, int kt, const void* keys_in, uint64_t n, int pay_bytes,
                 int digit_bits, void* ws, uint64_t ws_bytes, hipStream_t s,
                 bool with_hist = false);

// Run all scatter passes (single chunked pass when max-min < 2048, else the
// layout passes with constant digits skipped).
// Like sort_prepare for unsigned keys known to lie in [0, 2^key_bits): no
// host read, every digit below key_bits is sorted (graph-capturable).
int sort_prepare_known(SortCtx& c, int kt, const void* keys_in, uint64_t n, int pay_bytes,
                       int digit_bits, uint32_t key_bits, void* ws, uint64_t ws_bytes,
                       hipStream_t s, int shape = 0, bool small_tiles = false);
int sort_execute(SortCtx& ctx, const void* keys_in, void* keys_out, int want_keys,
                 const void* pay_in, void* pay_out, int iota, int descending);

// Single-pass only: digit = ordered(key) - min. Runs the column scan and
// leaves digit totals in ctx.total, then scatters (keys written only when
// keys_out is non-null).
int sort_single_pass(SortCtx& ctx, const void* keys_in, void* keys_out, const void* pay_in,
                     void* pay_out, int iota, int descending);

// Single-pass only: run the column scan so ctx.total holds the count of
// every digit (ordered(key) - min), without scattering.
int sort_digit_totals(SortCtx& ctx);

int sort_dispatch(int kt, const void* keys_in, void* keys_out, int want_keys, int pay_bytes,
                  const void* pay_in, void* pay_out, uint64_t n, int descending, int digit_bits,
                  void* ws, uint64_t ws_bytes, hipStream_t s);

}  // namespace jfx
