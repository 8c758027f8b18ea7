// Two-level preconditioner (coarse space): placeholder for the learned prolongation.
#include "../internal.h"

namespace drape {

struct TwoLevel {
  int coarse = 0;
};

void two_level_destroy(TwoLevel* t) { delete t; }

}  // namespace drape
