// ref_quadtree_shim.cpp — exposes the reference's CPU quadtree restatement
// (shader_validation/src/main.cpp, included in place from the reference tree,
// never copied) through two C entry points, so tests can check the oracle's
// neighbour-pair count against the reference's own quadtree walk.
//
// TEST INFRASTRUCTURE ONLY.  Fixed by the reference source: at most 1000
// agents (NUM_ENTITIES, main.cpp:152), world 29007.4609 x 16463.7656,
// collision radius 10 (main.cpp:153-156).
#include "main.cpp"

extern "C" {

// Inserts agents 0..n-1 at xy, then runs quad_tree_check_collisions for each
// agent in index order, as run_collision_detection_test_1 does
// (main.cpp:1137-1199).  Returns the pair counter (debugData[1]) or -1.
// in_pair (may be NULL) receives 1 for every agent quad_tree_collision
// recoloured (red in this restatement, main.cpp:741-745; blue in the shader),
// i.e. every agent found in at least one pair.
long long ref_quadtree_pairs(unsigned n, const float* xy, unsigned* in_pair) {
    if (n > NUM_ENTITIES) {
        return -1;
    }
    reset();
    init();
    for (unsigned i = 0; i < n; i++) {
        entities[i].pos = vec2(xy[2 * i], xy[2 * i + 1]);
        quad_tree_insert(i, 0, 1);
    }
    validate_entity_count(n);
    debugData[1] = 0;
    for (unsigned i = 0; i < n; i++) {
        entities[i].color = vec4(0, 1, 0, 1);
    }
    for (unsigned i = 0; i < n; i++) {
        quad_tree_check_collisions(i);
    }
    if (in_pair) {
        for (unsigned i = 0; i < n; i++) {
            in_pair[i] = entities[i].color.x == 1.0F ? 1U : 0U;
        }
    }
    return static_cast<long long>(debugData[1].load());
}

// The reference's known-answer test (asserts are live in this build).
int ref_known_answer_test(void) {
    run_collision_detection_test_1();
    return 0;
}

void ref_world(float* w, float* h, float* radius, unsigned* max_agents) {
    *w = pushConsts.worldSizeX;
    *h = pushConsts.worldSizeY;
    *radius = pushConsts.collisionRadius;
    *max_agents = static_cast<unsigned>(NUM_ENTITIES);
}
}
