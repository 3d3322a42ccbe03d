/*
 * nbrlist.cpp -- NeighborList (nbrlist.hpp in this directory) over the
 * device neighbour list of libmolmodel_amd.so.
 *
 * Reference behaviour followed: construction takes (n, cutoff, width)
 * without a box (nbrlist.cpp:9-17); set_boundary_conditions takes the box
 * and forces a build on the next coordinates_changed (nbrlist.cpp:84-122);
 * coordinates_changed rebuilds on the first call and when the skin test
 * fires (nbrlist.cpp:51-77).  The cell grid, the stencil, the same-cell
 * rule and the skin test run in mm_nlist_update; this file only moves rows
 * between the device and the arrays the accessors return.
 */
#include "nbrlist.hpp"
#include "bcond.hpp"
#include "molmodel_amd.h"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>

/* the reference's die() (util.c:11-26): message on stderr, exit status 1 */
static void fatal(const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  fprintf(stderr, "NeighborList: ");
  vfprintf(stderr, fmt, ap);
  fprintf(stderr, "\n");
  va_end(ap);
  exit(1);
}

NeighborList::NeighborList(int n_, double cutoff_, double width_)
  : n(n_), cutoff(cutoff_), width(width_), dev(0), fresh(1), cap(0),
    row_start(new int[n_ + 1]), cols(0), count(new int[n_]), row(new int *[n_])
{
  if (n <= 0) fatal("number of particles must be positive (got %d)", n);
  for (int i = 0; i <= n; i++) row_start[i] = 0;
  for (int i = 0; i < n; i++) { count[i] = 0; row[i] = 0; }
}

NeighborList::~NeighborList()
{
  if (dev) mm_nlist_destroy(dev);
  delete[] row_start;
  delete[] cols;
  delete[] count;
  delete[] row;
}

void NeighborList::set_boundary_conditions(const BoundaryConditions &bc)
{
  if (bc.type != cubic)
    fatal("only cubic boundary conditions run on the device (got type %d)", (int) bc.type);
  if (!(cutoff > 0)) fatal("cutoff must be positive (got %g)", cutoff);
  if (dev) {
    mm_nlist_destroy(dev);
    dev = 0;
  }
  /* a cube's inscribed-sphere diameter is its edge (bcond.cpp Cubic) */
  const double boxl = bc.min_diameter();
  if (mm_nlist_create(n, cutoff, width, boxl, &dev) != MM_OK)
    fatal("mm_nlist_create: %s", mm_last_error());
  fresh = 1;
}

void NeighborList::coordinates_changed(const CVec r)
{
  if (!dev) fatal("coordinates_changed before set_boundary_conditions");
  if (r.size() != n) fatal("expected %d positions, got %d", n, r.size());
  const int rebuilt = mm_nlist_update(dev, (const double *) r);
  if (rebuilt < 0) fatal("mm_nlist_update: %s", mm_last_error());
  if (rebuilt || fresh) fetch_rows();
  fresh = 0;
}

void NeighborList::fetch_rows()
{
  long np = mm_nlist_rows(dev, row_start, cols, cap);
  if (np < 0) fatal("mm_nlist_rows: %s", mm_last_error());
  if (np > cap) {
    delete[] cols;
    cap = np + np / 4 + 64;
    cols = new int[cap];
    np = mm_nlist_rows(dev, row_start, cols, cap);
    if (np < 0 || np > cap) fatal("mm_nlist_rows: %s", mm_last_error());
  }
  for (int i = 0; i < n; i++) {
    count[i] = row_start[i + 1] - row_start[i];
    row[i] = cols + row_start[i];
    std::sort(row[i], row[i] + count[i]);
  }
}

int * const * NeighborList::neighbors() const
{
  return row;
}

const int *NeighborList::number_of_neighbors() const
{
  return count;
}

const int *NeighborList::start(int i) const
{
  return row[i];
}

const int *NeighborList::end(int i) const
{
  return row[i] + count[i];
}

void NeighborList::show_self() const
{
  printf("Neighbor list (%d particles, cutoff %g, width %g):\n", n, cutoff, width);
  for (int i = 0; i < n; i++) {
    printf("%d:", i);
    for (int k = 0; k < count[i]; k++) printf(" %d", row[i][k]);
    printf("\n");
  }
  fflush(stdout);
}
