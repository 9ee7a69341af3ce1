// Test harness: a C++ caller of the reference's nine-function interface,
// written the way the reference app uses it, compiled with a plain host
// compiler against include/smokeSimulation.cuh and linked with
// libsmokesim.so -- no HIP headers, no GPU code of its own.
//
//  * A global volume object calls initializeVolume from its constructor, at
//    static-initialisation time, before main (project/main.cpp:87-88,
//    project/boundingBox.cpp:17-30).
//  * main: getGPUProperties (main.cpp:257), addSmokeSource at the grid centre
//    with r = 5 (:288-289), addObstacle(60,10,60, 0,0,0, 13) (:291), a first
//    update with dt = 0.01 (:293); then simulation ticks (:889-896) with the
//    source circling the centre (updateObjectPos, :609-614), the gravity and
//    buoyancy sliders written through their pointers (:832-833), and the
//    density copied out by each simulate (boundingBox.cpp:387-394); at the end
//    deleteVolume (boundingBox.cpp:397-402).
//
// Usage: app_driver <ticks> <out.bin>.  Writes, per tick (the first update
// included), the tick's inputs as six float32 -- source x, y, z, dt, gravity,
// buoyancy -- then the 80^3 density as float32, x fastest.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "smokeSimulation.cuh"

namespace {

struct Volume {
    int nx, ny, nz;
    std::vector<float> grid;
    Volume(int x, int y, int z) : nx(x), ny(y), nz(z), grid((size_t)x * y * z, 0.0f)
    {
        initializeVolume(grid.data(), (unsigned)nx, (unsigned)ny, (unsigned)nz);
    }
    void update(float dt) { simulate(grid.data(), dt); }
    void free_grid()
    {
        grid.clear();
        deleteVolume();
    }
};

Volume g_volume(80, 80, 80);  // constructed before main, as the app's bounding box

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <ticks> <out.bin>\n", argv[0]);
        return 2;
    }
    const int ticks = std::atoi(argv[1]);
    FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    auto dump = [&](float x, float y, float z, float dt) {
        const float head[6] = {x, y, z, dt, *getGravity(), *getBuoyancy()};
        if (std::fwrite(head, sizeof(float), 6, out) != 6 ||
            std::fwrite(g_volume.grid.data(), sizeof(float), g_volume.grid.size(), out) !=
                g_volume.grid.size())
            std::exit(3);
    };
    getGPUProperties();
    const float cx = 40.0f, cy = 40.0f, cz = 40.0f;
    const int source = addSmokeSource(cx, cy, cz, 5);
    const int obstacle = addObstacle(60, 10, 60, 0, 0, 0, 13);
    if (source != 0 || obstacle != 1) return 4;
    g_volume.update(0.01f);
    dump(cx, cy, cz, 0.01f);
    for (int k = 1; k < ticks; ++k) {
        const float t = 0.05f * (float)k;
        const float x = cx + 20.0f * std::cos(t), z = cz + 20.0f * std::sin(t);
        updateObjectPos(source, x, cy, z);
        if (k == 4) {
            *getGravity() = 7.5f;
            *getBuoyancy() = 4.0f;
        }
        const float dt = k % 3 == 0 ? 0.06f : 0.05f;
        g_volume.update(dt);
        dump(x, cy, z, dt);
    }
    std::fclose(out);
    g_volume.free_grid();
    return 0;
}
