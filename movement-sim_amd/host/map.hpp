// map.hpp — maps in the reference's JSON format for C++ hosts
// (sim::Map::load_from_file, src/sim/Map.cpp:28-150).
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "msim/types.h"

namespace msim_host {

struct Map {
    float width = 0.0f;   // maxDistLat
    float height = 0.0f;  // maxDistLong
    std::vector<msim_road> roads;
    std::vector<uint32_t> connections;
    uint32_t skipped = 0;  // zero-length roads dropped
};

// Map::load_from_file: nullptr when the file cannot be opened; throws
// std::runtime_error("Failed to parse map. '<field>' field missing.") like
// the reference on a missing field.  Zero-length roads are skipped (later
// road indices shift, the connection list is not renumbered).
std::shared_ptr<Map> load_map(const std::string& path);

// Appends one guard entry (road 0) when some group's top draw would read one
// past the connection list (random_move.comp:808 can return offset == count).
void add_guard_entry(Map& m);

}  // namespace msim_host
