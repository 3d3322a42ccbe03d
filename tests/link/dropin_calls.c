/*
 * dropin_calls.c -- link check for the drop-in boundary.
 *
 * Compiled against the reference's own headers (rspace.h, kspace.h, pppm.h,
 * cheby.h, pbin.h) and linked against libmolmodel_amd.so alone: every
 * function of those headers that the reference's C++ callers use
 * (eleclj.cpp, nbrlist.cpp, contact.cpp, ppot.cpp, msys.cpp, bcifit.cpp)
 * must resolve to the library.  Nothing is called: the table of addresses
 * only makes the linker look each symbol up.
 */
#include "rspace.h"
#include "kspace.h"
#include "pppm.h"
#include "cheby.h"
#include "pbin.h"

#include <stdio.h>

typedef void (*fn_t)(void);

static const fn_t used[] = {
  (fn_t)RSpaceChebyshev_new, (fn_t)RSpaceEwald, (fn_t)RSpaceEwaldPhi,
  (fn_t)RSpaceEwaldPhiAtGridpoints, (fn_t)RSpaceNonPeriodic, (fn_t)RSpaceNonPeriodicPhi,
  (fn_t)RSpaceNonPeriodicPhiAtGridpoints,
  (fn_t)kspace_new, (fn_t)kspace_delete, (fn_t)kspace_show, (fn_t)kspace_field,
  (fn_t)kspace_phi, (fn_t)kspace_number_of_wavevectors,
  (fn_t)p3m_new, (fn_t)p3m_copy, (fn_t)p3m_init, (fn_t)p3m_optimize, (fn_t)p3m_field,
  (fn_t)p3m_scale, (fn_t)p3m_delete,
  (fn_t)Chebyshev_init, (fn_t)Chebyshev_test,
  (fn_t)PairBinary_new, (fn_t)PairBinary_delete, (fn_t)PairBinary_add,
  (fn_t)PairBinary_quick_add, (fn_t)PairBinary_init, (fn_t)PairBinary_clear,
  (fn_t)PairBinary_number_of_elements, (fn_t)PairBinary_elements, (fn_t)PairBinary_start,
  (fn_t)PairBinary_end, (fn_t)PairBinary_exists, (fn_t)PairBinary_show,
};

int main(void)
{
  int n = 0;
  for (unsigned k = 0; k < sizeof used / sizeof used[0]; k++) n += used[k] != 0;
  printf("%d\n", n);
  return 0;
}
