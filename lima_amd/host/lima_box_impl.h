// lima_box_impl.h -- the owning container behind lima_box* (include/lima_gen.h), shared by the
// synthetic generators and the topology builder.  Arrays follow lima_box_desc (include/lima_hip.h).
#pragma once

#include <cstdint>
#include <vector>

struct lima_box {
    double box[3];
    int32_t nC = 0, nT = 0;
    std::vector<int32_t> cn, origo, key, coord;
    std::vector<float> vel, q, sig, eps, mass;
    std::vector<uint32_t> lut;
    std::vector<int32_t> tblock, tcoord;
    std::vector<float> tvel, tsig, teps, tmass;
    std::vector<int32_t> bond_ids, angle_ids, dih_ids, imp_ids, pair_ids;
    std::vector<float> bond_p, angle_p, dih_p, imp_p, pair_p;
};
