// The reference's nine-function simulation interface
// (project/smokeSimulation.cuh:4-17) as C entry points over one
// process-global simulation.  include/smokeSimulation.cuh wraps them with the
// reference's C++ names and its print-and-exit error behaviour.
//
// As in the reference, objects may be registered before or after the volume
// exists (its object list is a global vector, smokeSimulation.cu:54); ids are
// one counter shared by obstacles and sources (:88-104), and gravity/buoyancy
// are globals that the caller may write through the returned pointers
// (:28-36), read at every simulate().
#include <cstdio>
#include <mutex>
#include <vector>

#include "smoke.h"

namespace {

struct LegacyObject {
    int type;  // 0 obstacle, 1 source
    float x, y, z, vx, vy, vz, r;
};

std::mutex g_mu;
smoke_sim* g_sim = nullptr;
float g_gravity = -9.82f;  // smokeSimulation.cu:28
float g_buoyancy = 2.0f;   // smokeSimulation.cu:29
std::vector<LegacyObject> g_objects;

smoke_status push_object(const LegacyObject& o)
{
    int id = -1;
    if (o.type == 0) return smoke_add_obstacle(g_sim, o.x, o.y, o.z, o.vx, o.vy, o.vz, o.r, &id);
    return smoke_add_source(g_sim, o.x, o.y, o.z, o.r, &id);
}

int add(const LegacyObject& o)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_sim && push_object(o) != SMOKE_OK) return -1;
    g_objects.push_back(o);
    return (int)g_objects.size() - 1;
}

}  // namespace

extern "C" {

smoke_status smoke_legacy_device_info(void)
{
    char buf[256];
    smoke_status st = smoke_device_info(buf, sizeof buf);
    if (st == SMOKE_OK) printf("%s\n", buf);
    return st;
}

smoke_status smoke_legacy_initialize(float* grid, unsigned w, unsigned h, unsigned d)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_sim) {
        smoke_destroy(g_sim);
        g_sim = nullptr;
    }
    smoke_status st = smoke_create(w, h, d, grid, &g_sim);
    if (st != SMOKE_OK) return st;
    for (const LegacyObject& o : g_objects) {
        st = push_object(o);
        if (st != SMOKE_OK) return st;
    }
    return SMOKE_OK;
}

smoke_status smoke_legacy_delete(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    smoke_destroy(g_sim);
    g_sim = nullptr;
    return SMOKE_OK;
}

smoke_status smoke_legacy_simulate(float* grid, float dt)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_sim) return SMOKE_ERR_STATE;
    smoke_params p;
    smoke_status st = smoke_get_params(g_sim, &p);
    if (st != SMOKE_OK) return st;
    p.gravity = g_gravity;
    p.buoyancy = g_buoyancy;
    st = smoke_set_params(g_sim, &p);
    if (st != SMOKE_OK) return st;
    return smoke_step(g_sim, dt, grid);
}

int smoke_legacy_add_obstacle(float x, float y, float z, float vx, float vy, float vz, float r)
{
    return add(LegacyObject{0, x, y, z, vx, vy, vz, r});
}

int smoke_legacy_add_source(float x, float y, float z, float r)
{
    return add(LegacyObject{1, x, y, z, 0.0f, 0.0f, 0.0f, r});
}

smoke_status smoke_legacy_update_object_pos(int id, float x, float y, float z)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (id < 0 || id >= (int)g_objects.size()) return SMOKE_ERR_INVALID;
    g_objects[id].x = x;
    g_objects[id].y = y;
    g_objects[id].z = z;
    if (g_sim) return smoke_move_object(g_sim, id, x, y, z);
    return SMOKE_OK;
}

float* smoke_legacy_buoyancy_ptr(void) { return &g_buoyancy; }

float* smoke_legacy_gravity_ptr(void) { return &g_gravity; }

}  // extern "C"
