/*
 * nbrlist_driver.cpp -- uses the NeighborList of molmodel_amd/compat the
 * way ppot.cpp:54-60, 184-193 does: construct, set a cubic box, pass
 * positions, walk the rows.  Prints the pair count.  Without a GPU the
 * class stops the program with the library's error message.
 */
#include "nbrlist.hpp"
#include "bcond.hpp"

#include <cstdio>

int main()
{
  const int side = 8, n = side * side * side;
  const double boxl = 24.0, spacing = boxl / side;
  BoundaryConditions *bc = BoundaryConditions::new_cubic(boxl);
  CVec r(n);
  for (int i = 0; i < n; i++)
    r[i] = Cartesian((i % side + 0.25) * spacing, (i / side % side + 0.5) * spacing,
                     (i / (side * side) + 0.75) * spacing);
  const double rc = 0.45 * boxl;
  NeighborList *nlist = new NeighborList(n, rc, 0.499 * boxl - rc);
  nlist->set_boundary_conditions(*bc);
  nlist->coordinates_changed(r);
  long pairs = 0;
  for (int i = 0; i < n; i++) {
    const int nj = nlist->number_of_neighbors()[i];
    const int *j = nlist->neighbors()[i];
    for (int k = 0; k < nj; k++) pairs += j[k] != i;
  }
  printf("pairs %ld\n", pairs);
  delete nlist;
  delete bc;
  return 0;
}
