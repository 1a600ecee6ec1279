// C entry points over gs::ScratchState (gsplat-rs_amd/csrc/scratch_state.h)
// for tests/test_scratch_state.py, which builds this file with g++ and drives
// the state machine through random frame sequences against a word-level
// model of the scratch.
#include "scratch_state.h"

extern "C" {

void *ss_new() { return new gs::ScratchState(); }
void ss_free(void *s) { delete static_cast<gs::ScratchState *>(s); }
void ss_before(void *s, int lookback, unsigned long long lookback_words, unsigned long long *fill, int *skip_clear) {
  const gs::ScratchState::Plan p = static_cast<gs::ScratchState *>(s)->before(lookback != 0, lookback_words);
  *fill = p.fill;
  *skip_clear = p.skip_clear ? 1 : 0;
}
void ss_after(void *s, int lookback, unsigned long long rts_words) {
  static_cast<gs::ScratchState *>(s)->after(lookback != 0, rts_words);
}
void ss_reset(void *s) { static_cast<gs::ScratchState *>(s)->reset(); }

}  // extern "C"
