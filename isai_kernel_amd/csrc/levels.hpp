// levels.hpp -- host level-set analysis shared by ILU(0) and the triangular solves.
#pragma once

namespace isai {

int compute_levels(int nb, const int *rp, const int *ci, bool lower, int *level, int *nlevels);
// order[] = rows sorted by (level, row); level_ptr[nlevels+1] (may be null).
void level_order(int nb, const int *level, int nlevels, int *order, int *level_ptr);

}  // namespace isai
