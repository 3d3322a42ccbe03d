/*
 * eleclj_gpu.hpp -- a Potential (pot.hpp:6-28) that evaluates molmodel's
 * ElecLJ energy and forces on the GPU through libmolmodel_amd.so.
 *
 * It is written against ElecLJ's public surface only (eleclj.hpp:308-358):
 * the io parameters of a configured ElecLJ are copied at construction (its
 * public copy constructor), and init() derives sites, groups, exclusions
 * and per-site parameters from the MSys the way ElecLJ::init,
 * ElecLJ::types_changed and ElecLJ::init_sites do (eleclj.cpp:140-265,
 * 327-502, 643-771), through mm_derive_sites and mm_desc_resolve.
 *
 * Covered: cubic periodic boxes; atom sites with fixed charges and LJ
 * sigma/epsilon by type (general_type aliases included); Ewald or P3M with
 * real-space or ik differentiation; the LJ tail correction; box scaling; types
 * and lambda changes (types_changed(perturbed) + set_lambda).  Anything
 * else that ElecLJ accepts -- virtual (M, sp2, sp3) sites, bond charge
 * increments, 1-3 interactions, screening radii, 1-4 scaling, non-periodic
 * or non-cubic boxes, switching off the real-space, k-space or self
 * terms -- stops the program with a message in init().
 *
 * Built inside the reference's tree with its headers on the include path
 * (eleclj.hpp, msys.hpp, ...) and linked with libmolmodel_amd.so.
 */
#ifndef ELECLJ_GPU_H
#define ELECLJ_GPU_H

#include "eleclj.hpp"

#include <vector>

struct mm_engine;

class ElecLJGpu : public Potential
{
public:

  explicit ElecLJGpu(const ElecLJ &parameters);
  ~ElecLJGpu();

  void init(MSys *m);
  void add_to_energy_and_forces(double &u, Cartesian *f);
  void write() const;
  void scale(double s);
  void types_changed();
  void types_changed(const MSys &perturbed);
  void set_lambda(double lambda);

  double electrostatic_energy() const { return elec; }
  double lj_energy() const { return lj; }   /* pair LJ plus the tail correction */
  int number_of_sites() const { return nsite; }

private:

  ElecLJGpu(const ElecLJGpu &);
  ElecLJGpu &operator=(const ElecLJGpu &);

  const ElecLJ par;                  /* io parameters */
  struct mm_engine *eng;
  int nsite;
  std::vector<int> atom_of_site, mask;
  std::vector<double> pos, force;
  double elec, lj, lambda;

  void site_params(const MSys &m, std::vector<double> &q, std::vector<double> &sig,
                   std::vector<double> &eps, std::vector<int> &site_mask) const;
  void release();
};

#endif /* ELECLJ_GPU_H */
