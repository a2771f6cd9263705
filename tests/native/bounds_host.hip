// Host build of the bounds-check helper (CPU tests): built once with
// GFC_BOUNDS_CHECK (libboundshost.so) and once without (libboundsoff.so).
#include "../../gpufoam_amd/csrc/bounds.hpp"

extern "C" int bh_bound(int i, int n, unsigned* word, unsigned code) {
    return gfc::bound(i, n, word, code);
}
extern "C" int bh_checked(void) { return gfc::kBoundsChecked ? 1 : 0; }
