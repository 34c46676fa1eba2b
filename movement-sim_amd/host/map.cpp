// map.cpp — see map.hpp.  Field order and messages follow Map.cpp:28-150.
#include "map.hpp"

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "json_min.hpp"

namespace msim_host {

namespace {

const JsonValue& field(const JsonValue& o, const char* name) {
    if (!o.has(name)) throw std::runtime_error(std::string("Failed to parse map. '") + name + "' field missing.");
    return o.at(name);
}

float as_float(const JsonValue& v, const char* name) {
    if (v.kind != JsonValue::kNumber) {
        throw std::runtime_error(std::string("Failed to parse map. '") + name + "' is not a number.");
    }
    return static_cast<float>(v.num);
}

uint32_t as_unsigned(const JsonValue& v, const char* name) {
    if (v.kind != JsonValue::kNumber || !v.integral || v.num < 0.0 || v.num >= 4294967296.0) {
        throw std::runtime_error(std::string("Failed to parse map. '") + name +
                                 "' is not an unsigned 32-bit integer.");
    }
    return static_cast<uint32_t>(v.num);
}

}  // namespace

std::shared_ptr<Map> load_map(const std::string& path) {
    std::ifstream f(path);
    if (!f.is_open()) {
        std::fprintf(stderr, "Failed to open map from '%s'.\n", path.c_str());
        return nullptr;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    const JsonValue j = JsonReader(text).parse();
    auto m = std::make_shared<Map>();
    m->width = as_float(field(j, "maxDistLat"), "maxDistLat");
    m->height = as_float(field(j, "maxDistLong"), "maxDistLong");
    const JsonValue& roads = field(j, "roads");
    if (roads.kind != JsonValue::kArray) throw std::runtime_error("Failed to parse map. 'roads' is not an array.");
    for (const JsonValue& r : roads.arr) {
        const uint32_t i_s = as_unsigned(field(r, "connIndexStart"), "connIndexStart");
        const uint32_t c_s = as_unsigned(field(r, "connCountStart"), "connCountStart");
        const JsonValue& js = field(r, "start");
        const float sx = as_float(field(js, "distLat"), "distLat");
        const float sy = as_float(field(js, "distLong"), "distLong");
        const uint32_t i_e = as_unsigned(field(r, "connIndexEnd"), "connIndexEnd");
        const uint32_t c_e = as_unsigned(field(r, "connCountEnd"), "connCountEnd");
        const JsonValue& je = field(r, "end");
        const float ex = as_float(field(je, "distLat"), "distLat");
        const float ey = as_float(field(je, "distLong"), "distLong");
        if (sx == ex && sy == ey) {
            std::fprintf(stderr, "Zero length road detected. Skipping...\n");
            m->skipped++;
            continue;
        }
        msim_road road{};
        road.start = msim_coordinate{msim_vec2{sx, sy}, i_s, c_s};
        road.end = msim_coordinate{msim_vec2{ex, ey}, i_e, c_e};
        m->roads.push_back(road);
    }
    const JsonValue& conns = field(j, "connectionRoadIndexList");
    if (conns.kind != JsonValue::kArray) {
        throw std::runtime_error("Failed to parse map. 'connectionRoadIndexList' is not an array.");
    }
    m->connections.reserve(conns.arr.size());
    for (const JsonValue& c : conns.arr) m->connections.push_back(as_unsigned(c, "connectionRoadIndexList"));
    return m;
}

void add_guard_entry(Map& m) {
    uint64_t top = 0;
    bool any = false;
    for (const msim_road& r : m.roads) {
        for (const msim_coordinate* c : {&r.start, &r.end}) {
            uint64_t reach;
            if (c->connectedCount >= 3) {
                reach = static_cast<uint64_t>(c->connectedIndex) + c->connectedCount;
            } else if (c->connectedCount == 2) {
                reach = static_cast<uint64_t>(c->connectedIndex) + 1;
            } else {
                continue;
            }
            top = any ? std::max(top, reach) : reach;
            any = true;
        }
    }
    while (any && top >= m.connections.size()) m.connections.push_back(0u);
}

}  // namespace msim_host
