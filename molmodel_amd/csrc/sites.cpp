// sites.cpp -- sites, groups and excluded pairs from a bond graph, as
// ElecLJ::init_sites derives them (eleclj.cpp:643-771), for molecules whose
// sites are their atoms.  Host code only.
#include "../../include/molmodel_amd.h"
#include "engine.hpp"
#include <algorithm>
#include <utility>
#include <vector>

using namespace mmamd;

extern "C" int mm_derive_sites(int natom, const int *nbr_start, const int *nbr,
                               int *atom_of_site, int *group, int *ngroup, int *excl,
                               long maxexcl, long *nexcl, long *ncross) {
  if (natom <= 0 || !nbr_start || (!nbr && nbr_start[natom] > 0) || !atom_of_site || !group ||
      !ngroup || !nexcl || !ncross || maxexcl < 0 || (maxexcl > 0 && !excl)) {
    set_last_error("mm_derive_sites: bad arguments");
    return MM_ERR_ARG;
  }
  auto deg = [&](int a) { return nbr_start[a + 1] - nbr_start[a]; };
  for (int a = 0; a < natom; a++) {
    if (deg(a) < 0) {
      set_last_error("mm_derive_sites: nbr_start must not decrease");
      return MM_ERR_ARG;
    }
    for (int k = nbr_start[a]; k < nbr_start[a + 1]; k++) {
      const int b = nbr[k];
      if (b < 0 || b >= natom || b == a) {
        set_last_error("mm_derive_sites: neighbour index out of range or self-bond");
        return MM_ERR_ARG;
      }
      if (std::find(nbr + nbr_start[b], nbr + nbr_start[b + 1], a) == nbr + nbr_start[b + 1]) {
        set_last_error("mm_derive_sites: bonds must be listed from both atoms");
        return MM_ERR_ARG;
      }
      if (std::find(nbr + nbr_start[a], nbr + k, b) != nbr + k) {
        set_last_error("mm_derive_sites: a bond is listed twice");
        return MM_ERR_ARG;
      }
    }
  }
  // eleclj.cpp:653-677: a non-terminal (or isolated) atom opens a group and
  // takes its terminal neighbours in neighbour order; two terminal atoms
  // bonded to each other form one group, opened by the lower index
  std::vector<int> site_of_atom(natom, -1);
  int n = 0, ng = 0;
  auto place = [&](int a) {
    site_of_atom[a] = n;
    atom_of_site[n++] = a;
  };
  for (int a = 0; a < natom; a++) {
    if (deg(a) == 1) {
      const int b = nbr[nbr_start[a]];
      if (deg(b) == 1 && a < b) {
        group[2 * ng] = n;
        place(a);
        place(b);
        group[2 * ng + 1] = n;
        ng++;
      }
    } else {
      group[2 * ng] = n;
      place(a);
      for (int k = nbr_start[a]; k < nbr_start[a + 1]; k++)
        if (deg(nbr[k]) == 1) place(nbr[k]);
      group[2 * ng + 1] = n;
      ng++;
    }
  }
  // eleclj.cpp:678: with symmetric, duplicate-free bonds every atom is
  // placed exactly once (a terminal atom has one neighbour, hence one group)
  if (n != natom) {
    set_last_error("mm_derive_sites: internal error, sites and atoms differ in number");
    return MM_ERR_ARG;
  }
  *ngroup = ng;
  std::vector<int> group_of_site(natom);
  for (int g = 0; g < ng; g++)
    for (int s = group[2 * g]; s < group[2 * g + 1]; s++) group_of_site[s] = g;
  // eleclj.cpp:714-726: 1-2 pairs and pairs of neighbours of one site
  std::vector<std::pair<int, int>> pairs;
  auto add = [&](int i, int j) { pairs.emplace_back(std::min(i, j), std::max(i, j)); };
  for (int a = 0; a < natom; a++) {
    const int i = site_of_atom[a];
    for (int k = nbr_start[a]; k < nbr_start[a + 1]; k++) {
      add(i, site_of_atom[nbr[k]]);
      for (int l = k + 1; l < nbr_start[a + 1]; l++)
        if (nbr[l] != nbr[k]) add(site_of_atom[nbr[k]], site_of_atom[nbr[l]]);
    }
  }
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  long cross = 0;
  for (size_t k = 0; k < pairs.size(); k++) {
    if ((long)k < maxexcl) {
      excl[2 * k] = pairs[k].first;
      excl[2 * k + 1] = pairs[k].second;
    }
    cross += group_of_site[pairs[k].first] != group_of_site[pairs[k].second];
  }
  *nexcl = (long)pairs.size();
  *ncross = cross;
  return MM_OK;
}
