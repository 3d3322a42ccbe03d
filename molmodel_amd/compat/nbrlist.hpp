/*
 * nbrlist.hpp -- source drop-in for molmodel's NeighborList
 * (include/nbrlist.hpp:8-40, src/nbrlist.cpp), with the list built on the
 * GPU by libmolmodel_amd.so (mm_nlist_*).
 *
 * The public members are the reference's, with the same meaning: rows in
 * the PairBinary convention (each pair once, in the row of its canonical
 * owner, pbin.c:6), every pair within cutoff + width under the minimum
 * image plus every pair sharing a cell, rebuilt only when the largest and
 * second-largest displacement since the last build add up to more than
 * width (nbrlist.cpp:51-77).  Rows come back sorted by partner index (the
 * reference leaves them in insertion order); the pair set is the same.
 *
 * Compiled inside the reference's tree in place of src/nbrlist.cpp, with
 * this directory ahead of include/ so that callers (ppot.cpp, contact.cpp,
 * eleclj.cpp) pick this declaration up.  Only cubic boundary conditions run
 * on the device; any other type stops the program with a message, as the
 * reference does for conditions it does not handle.
 */
#ifndef NBRLIST_H
#define NBRLIST_H

#include "cvec.hpp"

class BoundaryConditions;
struct mm_nlist;

class NeighborList
{

public:

  NeighborList(int n, double cutoff, double width);
  ~NeighborList();

  void set_boundary_conditions(const BoundaryConditions &);
  void coordinates_changed(const CVec r);
  void show_self() const;

  int * const * neighbors() const; /* row i: the partners of i */
  const int *number_of_neighbors() const;
  const int *start(int i) const;
  const int *end(int i) const;

private:

  NeighborList(const NeighborList &);             /* owns a device list */
  NeighborList &operator=(const NeighborList &);

  const int n;
  const double cutoff, width;
  struct mm_nlist *dev;   /* device list of the current box; 0 before set_boundary_conditions */
  int fresh;              /* rows must be fetched on the next coordinates_changed */
  long cap;               /* capacity of cols */
  int *row_start;         /* n+1 offsets into cols */
  int *cols;              /* partners, row after row */
  int *count;             /* n row lengths */
  int **row;              /* n row pointers into cols */

  void fetch_rows();

};

#endif /* NBRLIST_H */
