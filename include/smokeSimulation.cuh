/*
 * smokeSimulation.cuh -- drop-in replacement for the reference's simulation
 * header (project/smokeSimulation.cuh:4-17 of RasmusAlmryd/smoke-simulation).
 *
 * The file name is kept so that the renderer's existing #include lines
 * (project/main.cpp:30, project/boundingBox.cpp:9) resolve unchanged.  It
 * holds no GPU API: each of the nine functions is an inline C++ wrapper over
 * the C ABI of libsmokesim.so (smoke.h), with the reference's names,
 * signatures and behaviour -- synchronous, one simulation per process, and
 * print-to-stderr-then-exit(-1) on any error (e.g. smokeSimulation.cu:131-135,
 * :815-818).
 */
#ifndef SMOKE_SIMULATION_CUH_COMPAT
#define SMOKE_SIMULATION_CUH_COMPAT

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "smoke.h"

namespace smoke_compat_detail {
inline void check(smoke_status st, const char* what)
{
    if (st != SMOKE_OK) {
        std::fprintf(stderr, "Error in %s: %s\n", what, smoke_last_error());
        std::exit(-1);
    }
}
}  // namespace smoke_compat_detail

/* smokeSimulation.cuh:4 -- prints the device description; exits if none. */
inline void getGPUProperties(void)
{
    smoke_compat_detail::check(smoke_legacy_device_info(), "getGPUProperties");
}

/* smokeSimulation.cuh:6 -- one time step; the new density (w*h*d floats) is
 * written into smoke_grid before returning. */
inline void simulate(float* smoke_grid, float dt)
{
    smoke_compat_detail::check(smoke_legacy_simulate(smoke_grid, dt), "simulate");
}

/* smokeSimulation.cuh:8 -- allocates the volume; smoke_grid is the initial
 * density (w*h*d floats, x fastest). */
inline void initializeVolume(float* smoke_grid, unsigned int width, unsigned int heigth,
                             unsigned int depth)
{
    smoke_compat_detail::check(smoke_legacy_initialize(smoke_grid, width, heigth, depth),
                               "initializeVolume");
}

/* smokeSimulation.cuh:9 */
inline void deleteVolume() { smoke_compat_detail::check(smoke_legacy_delete(), "deleteVolume"); }

/* smokeSimulation.cuh:12 -- returns the object id. */
inline int addObstacle(float x, float y, float z, float vx, float vy, float vz, float r)
{
    int id = smoke_legacy_add_obstacle(x, y, z, vx, vy, vz, r);
    if (id < 0) smoke_compat_detail::check(SMOKE_ERR_INVALID, "addObstacle");
    return id;
}

/* smokeSimulation.cuh:13 -- returns the object id. */
inline int addSmokeSource(float x, float y, float z, float r)
{
    int id = smoke_legacy_add_source(x, y, z, r);
    if (id < 0) smoke_compat_detail::check(SMOKE_ERR_INVALID, "addSmokeSource");
    return id;
}

/* smokeSimulation.cuh:14 */
inline void updateObjectPos(int id, float x, float y, float z)
{
    smoke_compat_detail::check(smoke_legacy_update_object_pos(id, x, y, z), "updateObjectPos");
}

/* smokeSimulation.cuh:16-17 -- pointers to the globals read by every
 * simulate(). */
inline float* getBuoyancy() { return smoke_legacy_buoyancy_ptr(); }
inline float* getGravity() { return smoke_legacy_gravity_ptr(); }

#endif /* SMOKE_SIMULATION_CUH_COMPAT */
