// scratch_state.h -- bookkeeping of a radix sort's scratch words across the
// frames of a pipeline (host only, no device code).
//
// A look-back sort needs every word it reads to start at zero: digit
// histograms are accumulated with atomics, tile tickets count from zero, and
// a status word whose flag bits are set is taken as published. A pipeline
// therefore clears a look-back sort's words with a later kernel of the same
// frame, so the next frame's sort can skip its own clear. A reduce-then-scan
// sort writes the same scratch differently: digit counts, then flagged scan
// status words beyond them, which no look-back clean-up covers. A look-back
// sort that followed one could read such a stale flagged word as a
// predecessor's published prefix whenever the real predecessor had not yet
// published, a race that scrambles the order. So every word a
// reduce-then-scan sort may have written is zeroed (by a fill kernel on the
// pipeline's stream) before the next look-back sort.
#pragma once

#include <algorithm>
#include <cstddef>

namespace gs {

class ScratchState {
 public:
  struct Plan {
    size_t fill = 0;          // words [0, fill) to zero before the sort (0: none)
    bool skip_clear = false;  // the look-back passes may skip their own clear
  };
  // Before a sort in look-back mode (lookback) or reduce-then-scan mode,
  // whose look-back words are [0, lookback_words).
  Plan before(bool lookback, size_t lookback_words) const {
    Plan p;
    if (!lookback) return p;
    if (rts_dirty_) p.fill = std::max(rts_dirty_, lookback_words);
    p.skip_clear = true;
    return p;
  }
  // After the frame. A look-back sort's words [0, lookback_words) must have
  // been cleared by a later kernel of the frame; a reduce-then-scan sort
  // leaves [0, rts_words) dirty.
  void after(bool lookback, size_t rts_words) {
    if (lookback)
      rts_dirty_ = 0;
    else
      rts_dirty_ = std::max(rts_dirty_, rts_words);
  }
  // The scratch was (re)allocated and zeroed.
  void reset() { rts_dirty_ = 0; }
  size_t rts_dirty() const { return rts_dirty_; }

 private:
  size_t rts_dirty_ = 0;  // words a reduce-then-scan sort may have left non-zero
};

}  // namespace gs
