// adg_regrid.hpp -- host side of the tagging-driven regrid: new box layouts
// of the fine levels from the refinement tags, as AmrCore::regrid(0, t) ->
// MakeNewGrids does it for the reference's Mesh (Core/Mesh.h:128-196,
// blocking factor 4 at Mesh.h:146; the regrid points of AmrDG::evolve,
// AmrDG.h:936-986).  The clustering is Berger-Rigoutsos (Berger & Rigoutsos,
// IEEE Trans. Syst. Man Cybern. 21(5), 1991): a cluster whose fill ratio is
// below grid_eff is cut at the signature hole nearest its centre, else at
// the strongest sign change of the signature's second difference, else in
// the middle.
#pragma once

#include <array>
#include <vector>

namespace adg {

struct IBox {
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
};

struct AmrSettings {
    int max_level = 0;          // index of the finest level allowed
    int blocking_factor = 4;    // Mesh::init, SetBlockingFactor(4)
    int max_grid_size = 128;    // AMReX default: 128 in 2D, 32 in 3D
    double grid_eff = 0.7;      // AMReX default
    int n_error_buf = 1;        // AMReX default
    int n_proper = 1;           // AMReX default
};

// Tagged cells of one level, in that level's index space.
using TagList = std::vector<std::array<int, 3>>;

// New boxes of levels 1..new_finest (index l-1 of the result holds level l)
// from the tags of levels 0..min(old_finest, max_level-1) (tags[l]) and the
// current boxes of those levels (old[l], for the buffer region of the tag
// arrays).  Level 0 is never changed (it covers the domain).
std::vector<std::vector<IBox>> make_new_grids(int D, const int* dom0, const int* periodic, const AmrSettings& s,
                                              const std::vector<std::vector<IBox>>& old,
                                              const std::vector<TagList>& tags);

// Berger-Rigoutsos clustering of tagged points (any index space): boxes that
// cover every point, each with fill ratio >= eff or no admissible cut.
std::vector<IBox> cluster_tags(int D, const TagList& pts, double eff);

}  // namespace adg
