// Smoke time-step ORACLE -- test infrastructure only.
//
// A CPU restatement of the reference simulation step, written from
// project/smokeSimulation.cu (simulate(), :774-819) and its kernels.  Only the
// test suite, the smoke check and the CPU-baseline leg of the benchmark load
// this library; it is the checker and is never part of the product path.
//
// The template parameter Real selects the arithmetic:
//   Real = double : the FP64 oracle (every float operation of the reference is
//                   carried out in double; float-valued inputs are widened).
//   Real = float  : the FP32 restatement.  It keeps the reference's two
//                   double-promoted expressions (the SOR factor product,
//                   smokeSimulation.cu:384 with OVER_RELAXATION a double
//                   literal at :38, and the density back-trace, :628-630) and
//                   otherwise evaluates every expression in the reference's
//                   operand order.  Built with -ffp-contract=off it is the
//                   bit-for-bit spec the FP32 GPU kernels are held to.
//
// Layout follows the reference exactly: density and mask are nx*ny*nz with
// c = x + y*nx + z*nx*ny (:267); each velocity component is
// (nx+1)*(ny+1)*(nz+1) with f = x + y*Sx + z*Sx*Sy, S = n+1 (:146-147).  Both
// ping-pong halves of every field are kept, because entries that a step does
// not write keep the value from two steps earlier and are read again.
//
// Extensions beyond the reference (each defaults to the reference behaviour):
//   * a caller-supplied base solid mask (the reference only has the y = 0
//     floor, :200-207) and an obstacle-combination mode (reference: the last
//     obstacle overwrites the interior mask, :295-311);
//   * a density gate for buoyancy (reference: none);
//   * the clamp constant M (reference MAX_VELOCITY_PER_STEP = 3, :39);
//   * two pressure-form projections on the same linear system as the
//     reference's in-place red-black SOR: red-black SOR and (damped) Jacobi on
//     P with L_s P = -div(u*), followed by a masked gradient subtraction.
//
// Pinning status of each function (how the tests in tests/ hold it):
//   sphere rasteriser / mask / flags ... brute force on small grids, floor
//                                        plane, last-obstacle-wins (bit-exact)
//   sources ............................ closed form
//   buoyancy / clamp ................... closed forms incl. the clamp boundary
//   sampler ............................ affine exactness, wall clamping,
//                                        max principle
//   velocity / density advection ....... zero-velocity identity, integer-shift
//                                        exactness, stale-value retention
//   projection (velocity form) ......... divergence-free fixed point, colour
//                                        disjointness, equality with the
//                                        pressure form, convergence to a dense
//                                        direct solve
//   pressure-form Jacobi / RB-SOR ...... dense direct solve at convergence,
//                                        agreement with each other, one hand
//                                        sweep from P = 0
// No reference-held golden vector exists for this path (the reference ships
// no tests or fixtures), so none of these is pinned against reference output.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

enum Projection { PROJ_REFERENCE = 0, PROJ_RBSOR = 1, PROJ_JACOBI = 2 };
enum ObstacleMode { OBST_LAST_WINS = 0, OBST_UNION = 1 };

enum Stage {
    STAGE_SCENE = 0,        // swap halves, sources, solid mask
    STAGE_FORCE = 1,        // buoyancy + speed clamp on [now]
    STAGE_PROJECT = 2,      // pressure projection on [now]
    STAGE_ADVECT_VEL = 3,   // u,v,w [now] -> [past]
    STAGE_ADVECT_DENS = 4,  // density [now] -> [past]
};

// Cell flag bits: self fluid, the six face neighbours fluid, and "active"
// (interior, fluid, at least one fluid neighbour: the cells the projection
// relaxes, smokeSimulation.cu:364-377).
enum : uint8_t {
    F_SELF = 1, F_XM = 2, F_XP = 4, F_YM = 8, F_YP = 16, F_ZM = 32, F_ZP = 64,
    F_ACTIVE = 128
};
constexpr uint8_t F_NBRS = F_XM | F_XP | F_YM | F_YP | F_ZM | F_ZP;
// Vorticity confinement: no force where |grad |omega|| <= this (cells/s per
// cell^2), i.e. where the gradient is rounding noise and has no direction.
constexpr float kVortMinGrad = 1e-6f;

struct Params {
    float gravity;        // smokeSimulation.cu:28 (-9.82)
    float buoyancy;       // smokeSimulation.cu:29 (2.0)
    float density_gate;   // <= 0: no gate (reference)
    float max_velocity;   // smokeSimulation.cu:39 (3.0)
    double sor_omega;     // smokeSimulation.cu:38 (1.9)
    double jacobi_omega;  // 1.0 = plain Jacobi
    int projection;       // Projection
    int iterations;       // smokeSimulation.cu:797 (30)
    int obstacle_mode;    // ObstacleMode
    int warm_start;       // pressure forms: start from the previous P
    float vorticity;      // vorticity-confinement strength eps; 0 = off (reference)
    int obstacle_velocity;  // 1: moving-obstacle velocity boundary condition; 0 = off (reference)
};

struct Sphere {
    int id;
    float pos[3];
    float vel[3];
    float r;
    int type;  // 0 obstacle, 1 smoke source (smokeSimulation.cu:45-52)
};

Params default_params()
{
    Params p;
    p.gravity = -9.82f;
    p.buoyancy = 2.0f;
    p.density_gate = 0.0f;
    p.max_velocity = 3.0f;
    p.sor_omega = 1.9;
    p.jacobi_omega = 1.0;
    p.projection = PROJ_REFERENCE;
    p.iterations = 30;
    p.obstacle_mode = OBST_LAST_WINS;
    p.warm_start = 0;
    p.vorticity = 0.0f;
    p.obstacle_velocity = 0;
    return p;
}

struct SimBase {
    virtual ~SimBase() {}
    int nx = 0, ny = 0, nz = 0;
    long Sx = 0, Sy = 0, Sz = 0;
    long N = 0, NS = 0;
    std::vector<uint8_t> base, s, flags;
    // Per cell: id of the obstacle whose sphere made the cell solid (-1: none).
    std::vector<int> owner;
    std::vector<Sphere> objects;
    int now = 1;  // indexNow, smokeSimulation.cu:707
    Params prm = default_params();

    long c(int x, int y, int z) const { return x + (long)y * nx + (long)z * nx * ny; }
    long f(int x, int y, int z) const { return x + y * Sx + z * Sx * Sy; }
    bool interior(int x, int y, int z) const
    {
        return x >= 1 && x <= nx - 2 && y >= 1 && y <= ny - 2 && z >= 1 && z <= nz - 2;
    }

    virtual void set_density(int half, const double* in) = 0;
    virtual void get_density(int half, double* out) const = 0;
    virtual void set_velocity(int comp, int half, const double* in) = 0;
    virtual void get_velocity(int comp, int half, double* out) const = 0;
    virtual void get_pressure(double* out) const = 0;
    virtual void get_rhs(double* out) const = 0;
    virtual void run_stage(int stage, float dt) = 0;

    // Sphere membership test of fillSmoke/fillObstacle (smokeSimulation.cu:265,
    // :304): integer cell index against the float centre, squared in float.
    static bool inside(const Sphere& o, int x, int y, int z)
    {
        float dx = (float)x - o.pos[0];
        float dy = (float)y - o.pos[1];
        float dz = (float)z - o.pos[2];
        float dist = dx * dx + dy * dy + dz * dz;
        return dist < o.r * o.r;
    }

    // Solid mask from the base mask and the obstacle list
    // (fillObstacle, smokeSimulation.cu:289-313): only interior cells are
    // touched; with no obstacles the mask is the base mask.
    void rebuild_mask()
    {
        std::vector<const Sphere*> obst;
        for (const Sphere& o : objects)
            if (o.type == 0) obst.push_back(&o);
#pragma omp parallel for schedule(static)
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    long i = c(x, y, z);
                    uint8_t v = base[i];
                    int who = -1;
                    if (!obst.empty() && interior(x, y, z)) {
                        bool solid;
                        if (prm.obstacle_mode == OBST_LAST_WINS) {
                            solid = inside(*obst.back(), x, y, z);
                            if (solid) who = obst.back()->id;
                        } else {
                            solid = false;
                            for (const Sphere* o : obst)
                                if (inside(*o, x, y, z)) {
                                    solid = true;
                                    who = o->id;  // the last obstacle containing the cell
                                }
                        }
                        if (solid) v = 0;
                    }
                    s[i] = v;
                    owner[i] = who;
                }
        // Flags: neighbour bits exist only where the neighbour exists.
#pragma omp parallel for schedule(static)
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    long i = c(x, y, z);
                    uint8_t fl = s[i] ? F_SELF : 0;
                    if (x > 0 && s[i - 1]) fl |= F_XM;
                    if (x < nx - 1 && s[i + 1]) fl |= F_XP;
                    if (y > 0 && s[i - nx]) fl |= F_YM;
                    if (y < ny - 1 && s[i + nx]) fl |= F_YP;
                    if (z > 0 && s[i - (long)nx * ny]) fl |= F_ZM;
                    if (z < nz - 1 && s[i + (long)nx * ny]) fl |= F_ZP;
                    if (interior(x, y, z) && (fl & F_SELF) && (fl & F_NBRS)) fl |= F_ACTIVE;
                    flags[i] = fl;
                }
    }
};

template <class Real>
struct Sim : SimBase {
    std::vector<Real> rho[2], u[2], v[2], w[2];
    std::vector<Real> P, Q, d;

    Sim(int nx_, int ny_, int nz_)
    {
        nx = nx_; ny = ny_; nz = nz_;
        Sx = nx + 1; Sy = ny + 1; Sz = nz + 1;
        N = (long)nx * ny * nz;
        NS = Sx * Sy * Sz;
        for (int h = 0; h < 2; ++h) {
            // Half [1] is uninitialised device memory in the reference
            // (smokeSimulation.cu:130-189); it is defined as zero here.
            rho[h].assign(N, Real(0));
            u[h].assign(NS, Real(0));
            v[h].assign(NS, Real(0));
            w[h].assign(NS, Real(0));
        }
        P.assign(N, Real(0));
        Q.assign(N, Real(0));
        d.assign(N, Real(0));
        // Initial mask: fluid everywhere except the y = 0 floor
        // (smokeSimulation.cu:200-207).
        base.assign(N, 1);
        for (int z = 0; z < nz; ++z)
            for (int x = 0; x < nx; ++x) base[c(x, 0, z)] = 0;
        s = base;
        flags.assign(N, 0);
        owner.assign(N, -1);
        rebuild_mask();
    }

    std::vector<Real>& comp(int k, int h) { return k == 0 ? u[h] : (k == 1 ? v[h] : w[h]); }
    const std::vector<Real>& comp(int k, int h) const { return k == 0 ? u[h] : (k == 1 ? v[h] : w[h]); }

    void set_density(int half, const double* in) override
    {
        for (long i = 0; i < N; ++i) rho[half][i] = (Real)in[i];
    }
    void get_density(int half, double* out) const override
    {
        for (long i = 0; i < N; ++i) out[i] = (double)rho[half][i];
    }
    void set_velocity(int k, int half, const double* in) override
    {
        std::vector<Real>& a = comp(k, half);
        for (long i = 0; i < NS; ++i) a[i] = (Real)in[i];
    }
    void get_velocity(int k, int half, double* out) const override
    {
        const std::vector<Real>& a = comp(k, half);
        for (long i = 0; i < NS; ++i) out[i] = (double)a[i];
    }
    void get_pressure(double* out) const override
    {
        for (long i = 0; i < N; ++i) out[i] = (double)P[i];
    }
    void get_rhs(double* out) const override
    {
        for (long i = 0; i < N; ++i) out[i] = (double)d[i];
    }

    // ---------------------------------------------------------------- stages

    // smokeSimulation.cu:777-779 (swap) and drawObjects :714-771.
    void stage_scene()
    {
        now = 1 - now;
        // fillSmoke (smokeSimulation.cu:251-273): interior cells inside any
        // source are set to 1.0 in BOTH halves (the launch passes dev_smoke[0]
        // and dev_smoke[1], :739).
        std::vector<const Sphere*> src;
        for (const Sphere& o : objects)
            if (o.type == 1) src.push_back(&o);
        if (!src.empty()) {
#pragma omp parallel for schedule(static)
            for (int z = 1; z < nz - 1; ++z)
                for (int y = 1; y < ny - 1; ++y)
                    for (int x = 1; x < nx - 1; ++x)
                        for (const Sphere* o : src)
                            if (inside(*o, x, y, z)) {
                                rho[0][c(x, y, z)] = Real(1);
                                rho[1][c(x, y, z)] = Real(1);
                            }
        }
        rebuild_mask();
        if (prm.obstacle_velocity) obstacle_velocity_bc();
    }

    // Moving-obstacle velocity boundary condition, an extension: the
    // reference stores each obstacle's velocity (smokeSimulation.cu:45-52)
    // but never uses it (:299-301).  With it on, after the mask is built, every
    // face of [now] between two cells at least one of which an obstacle made
    // solid takes that obstacle's velocity component (the high-side cell's
    // obstacle when both are).  The projection leaves faces next to solid
    // cells alone and the advection writes only fluid-fluid faces, so the
    // obstacle's motion enters the divergence the projection removes.
    void obstacle_velocity_bc()
    {
        const long step[3] = {1, nx, (long)nx * ny};
#pragma omp parallel for schedule(static)
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    const long i = c(x, y, z);
                    const int pos[3] = {x, y, z};
                    for (int k = 0; k < 3; ++k) {
                        if (pos[k] < 1) continue;  // the low cell is outside the grid
                        const int who = owner[i] >= 0 ? owner[i] : owner[i - step[k]];
                        if (who < 0) continue;
                        comp(k, now)[f(x, y, z)] = (Real)objects[who].vel[k];
                    }
                }
    }

    // integrate (smokeSimulation.cu:315-329) then velocityConfinement
    // (:331-352), both on half [now].  Both are pointwise per staggered index,
    // so running them back to back per index equals the two launches.
    void stage_force(float dt_f)
    {
        const Real dt = (Real)dt_f;
        const Real g = (Real)prm.gravity;
        const Real alpha = (Real)prm.buoyancy;
        const Real M = (Real)prm.max_velocity;
        const Real M2 = M * M;
        const bool gate_on = prm.density_gate > 0.0f;
        const Real gate = (Real)prm.density_gate;
        std::vector<Real>& U = u[now];
        std::vector<Real>& V = v[now];
        std::vector<Real>& W = w[now];
        const std::vector<Real>& R = rho[now];
        // Buoyancy: x in [0,nx-1], y in [1,ny-1], z in [0,nz-1], both cells of
        // the v-face fluid; density of the upper cell only.
#pragma omp parallel for schedule(static)
        for (int z = 0; z < nz; ++z)
            for (int y = 1; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    if (s[c(x, y, z)] != 0 && s[c(x, y - 1, z)] != 0) {
                        Real r = R[c(x, y, z)];
                        if (gate_on && !(r > gate)) continue;
                        Real b = alpha * (r - Real(0));
                        long fi = f(x, y, z);
                        V[fi] += r * g * dt + b * dt;
                    }
                }
        // Clamp: x,y,z in [1,n-1], the three components at the SAME index.
#pragma omp parallel for schedule(static)
        for (int z = 1; z < nz; ++z)
            for (int y = 1; y < ny; ++y)
                for (int x = 1; x < nx; ++x) {
                    long fi = f(x, y, z);
                    Real ut = U[fi], vt = V[fi], wt = W[fi];
                    Real L = ut * ut + vt * vt + wt * wt;
                    if (L * dt > M2) {
                        U[fi] = ut * (M2 / (L * dt));
                        V[fi] = vt * (M2 / (L * dt));
                        W[fi] = wt * (M2 / (L * dt));
                    }
                }
        if (prm.vorticity > 0.0f) vorticity_confinement(dt);
    }

    // Vorticity confinement (listed as future work, README.md:33; no reference
    // code, so this defines it): with cell-centred velocities c = mean of the
    // two faces of a cell, omega = curl c by central differences on interior
    // fluid cells (0 elsewhere), eta = grad |omega| by central differences,
    // N = eta/|eta| (no force where |eta| <= kVortMinGrad) and F = eps * (N x omega) on
    // interior fluid cells; every face between two fluid cells gains
    // dt * (mean of its two cells' F).  Runs after the clamp on [now].
    Real ucen(const std::vector<Real>& A, int x, int y, int z, int axis) const
    {
        const long a = f(x, y, z);
        const long b = axis == 0 ? f(x + 1, y, z) : (axis == 1 ? f(x, y + 1, z) : f(x, y, z + 1));
        return (A[a] + A[b]) * Real(0.5);
    }
    void curl_at(int x, int y, int z, Real& ox, Real& oy, Real& oz) const
    {
        const std::vector<Real>& U = u[now];
        const std::vector<Real>& V = v[now];
        const std::vector<Real>& W = w[now];
        const Real h = Real(0.5);
        ox = (ucen(W, x, y + 1, z, 2) - ucen(W, x, y - 1, z, 2)) * h -
             (ucen(V, x, y, z + 1, 1) - ucen(V, x, y, z - 1, 1)) * h;
        oy = (ucen(U, x, y, z + 1, 0) - ucen(U, x, y, z - 1, 0)) * h -
             (ucen(W, x + 1, y, z, 2) - ucen(W, x - 1, y, z, 2)) * h;
        oz = (ucen(V, x + 1, y, z, 1) - ucen(V, x - 1, y, z, 1)) * h -
             (ucen(U, x, y + 1, z, 0) - ucen(U, x, y - 1, z, 0)) * h;
    }
    void vorticity_confinement(Real dt)
    {
        const Real eps = (Real)prm.vorticity;
        std::vector<Real> mag(N, Real(0)), F[3];
        for (auto& a : F) a.assign(N, Real(0));
#pragma omp parallel for schedule(static)
        for (int z = 1; z < nz - 1; ++z)
            for (int y = 1; y < ny - 1; ++y)
                for (int x = 1; x < nx - 1; ++x) {
                    if (!s[c(x, y, z)]) continue;
                    Real ox, oy, oz;
                    curl_at(x, y, z, ox, oy, oz);
                    mag[c(x, y, z)] = std::sqrt(ox * ox + oy * oy + oz * oz);
                }
#pragma omp parallel for schedule(static)
        for (int z = 1; z < nz - 1; ++z)
            for (int y = 1; y < ny - 1; ++y)
                for (int x = 1; x < nx - 1; ++x) {
                    const long i = c(x, y, z);
                    if (!s[i]) continue;
                    const Real h = Real(0.5);
                    const Real ex = (mag[i + 1] - mag[i - 1]) * h;
                    const Real ey = (mag[i + nx] - mag[i - nx]) * h;
                    const Real ez = (mag[i + (long)nx * ny] - mag[i - (long)nx * ny]) * h;
                    const Real len = std::sqrt(ex * ex + ey * ey + ez * ez);
                    // |grad|omega|| at rounding-noise level (a uniform vorticity
                    // gives ~1e-9 in FP32) has no direction: no force there.
                    if (!(len > Real(kVortMinGrad))) continue;
                    const Real nxv = ex / len, nyv = ey / len, nzv = ez / len;
                    Real ox, oy, oz;
                    curl_at(x, y, z, ox, oy, oz);
                    F[0][i] = eps * (nyv * oz - nzv * oy);
                    F[1][i] = eps * (nzv * ox - nxv * oz);
                    F[2][i] = eps * (nxv * oy - nyv * ox);
                }
        std::vector<Real>* vel[3] = {&u[now], &v[now], &w[now]};
        const long step[3] = {1, nx, (long)nx * ny};
#pragma omp parallel for schedule(static)
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    const long i = c(x, y, z);
                    const int pos[3] = {x, y, z};
                    for (int k = 0; k < 3; ++k) {
                        if (pos[k] < 1 || !s[i] || !s[i - step[k]]) continue;  // face between two fluid cells
                        (*vel[k])[f(x, y, z)] += dt * ((F[k][i - step[k]] + F[k][i]) * Real(0.5));
                    }
                }
    }

    int acc_of(uint8_t fl) const { return __builtin_popcount(fl & F_NBRS); }

    // Divergence of cell (x,y,z) in the reference's operand order
    // (smokeSimulation.cu:379-381).
    Real divergence(int x, int y, int z) const
    {
        const std::vector<Real>& U = u[now];
        const std::vector<Real>& V = v[now];
        const std::vector<Real>& W = w[now];
        return -U[f(x, y, z)] + U[f(x + 1, y, z)] + -V[f(x, y, z)] + V[f(x, y + 1, z)] +
               -W[f(x, y, z)] + W[f(x, y, z + 1)];
    }

    // Literal red-black SOR on face velocities (smokeSimulation.cu:356-394,
    // driven 'iterations' times per colour pair at :797-801).
    void project_reference()
    {
        std::vector<Real>& U = u[now];
        std::vector<Real>& V = v[now];
        std::vector<Real>& W = w[now];
        for (int it = 0; it < prm.iterations; ++it)
            for (int o = 0; o < 2; ++o) {
#pragma omp parallel for schedule(static)
                for (int z = 1; z < nz - 1; ++z)
                    for (int y = 1; y < ny - 1; ++y)
                        for (int x = 1; x < nx - 1; ++x) {
                            if (((x + y + z + o) & 1) != 0) continue;
                            long ci = c(x, y, z);
                            if (s[ci] == 0) continue;
                            int sx1 = s[ci + 1], sx0 = s[ci - 1];
                            int sy1 = s[ci + nx], sy0 = s[ci - nx];
                            int sz1 = s[ci + (long)nx * ny], sz0 = s[ci - (long)nx * ny];
                            int acc = sx0 + sx1 + sy0 + sy1 + sz0 + sz1;
                            if (acc == 0) continue;
                            Real dv = divergence(x, y, z);
                            // float / int -> float, then times a double literal,
                            // rounded back to float on assignment.
                            Real p = (Real)((double)(-dv / (Real)acc) * prm.sor_omega);
                            U[f(x, y, z)] -= p * (Real)sx0;
                            U[f(x + 1, y, z)] += p * (Real)sx1;
                            V[f(x, y, z)] -= p * (Real)sy0;
                            V[f(x, y + 1, z)] += p * (Real)sy1;
                            W[f(x, y, z)] -= p * (Real)sz0;
                            W[f(x, y, z + 1)] += p * (Real)sz1;
                        }
            }
    }

    // Neighbour sum of the pressure form at an active (hence interior) cell:
    // all six neighbours, left to right in the fixed order x-, x+, y-, y+, z-,
    // z+.  P is exactly +0 on every cell that is not active (never relaxed,
    // zeroed for warm starts), so this is the sum over the fluid neighbours
    // s_n P_n of the system acc_i P_i - sum_n s_n P_n = -d_i.
    Real nbr_sum(const std::vector<Real>& A, long ci, uint8_t fl) const
    {
        (void)fl;
        const long pl = (long)nx * ny;
        Real sum = A[ci - 1] + A[ci + 1];
        sum = sum + A[ci - nx];
        sum = sum + A[ci + nx];
        sum = sum + A[ci - pl];
        sum = sum + A[ci + pl];
        return sum;
    }

    // Diagonal inverse 1/acc, rounded to Real (the pressure forms multiply by
    // it instead of dividing).
    static Real inv_acc(int acc) { return Real(1) / (Real)acc; }

    // Pressure form of the same projection (derivation: each face of the
    // reference's in-place update equals u* - s_a s_b (P_b - P_a) with P the
    // accumulated per-cell corrections, so the colour sweep is SOR on
    // acc_i P_i - sum_n s_n P_n = -div_i(u*)).  The update of a cell is
    // G = (sum_n s_n P_n - d_i) * (1/acc_i), evaluated as one fused
    // multiply-add: G = fma(sum, inv, 0 - d*inv), inv = RN(1/acc).
    void project_pressure()
    {
        // Right-hand side: d = div(u*) on active cells.
#pragma omp parallel for schedule(static)
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    long ci = c(x, y, z);
                    d[ci] = (flags[ci] & F_ACTIVE) ? divergence(x, y, z) : Real(0);
                }
        // Initial guess.
        for (long i = 0; i < N; ++i) {
            if (!prm.warm_start || !(flags[i] & F_ACTIVE)) P[i] = Real(0);
            Q[i] = Real(0);
        }
        if (prm.projection == PROJ_JACOBI) {
            const bool plain = prm.jacobi_omega == 1.0;
            const Real om = (Real)prm.jacobi_omega;
            for (int it = 0; it < prm.iterations; ++it) {
                std::vector<Real>& A = P;
                std::vector<Real>& B = Q;
#pragma omp parallel for schedule(static)
                for (int z = 1; z < nz - 1; ++z)
                    for (int y = 1; y < ny - 1; ++y)
                        for (int x = 1; x < nx - 1; ++x) {
                            long ci = c(x, y, z);
                            uint8_t fl = flags[ci];
                            if (!(fl & F_ACTIVE)) { B[ci] = Real(0); continue; }
                            const Real inv = inv_acc(acc_of(fl));
                            const Real G = std::fma(nbr_sum(A, ci, fl), inv, Real(0) - d[ci] * inv);
                            B[ci] = plain ? G : A[ci] + om * (G - A[ci]);
                        }
                P.swap(Q);
            }
        } else {
            const Real om = (Real)prm.sor_omega;
            for (int it = 0; it < prm.iterations; ++it)
                for (int o = 0; o < 2; ++o) {
#pragma omp parallel for schedule(static)
                    for (int z = 1; z < nz - 1; ++z)
                        for (int y = 1; y < ny - 1; ++y)
                            for (int x = 1; x < nx - 1; ++x) {
                                if (((x + y + z + o) & 1) != 0) continue;
                                long ci = c(x, y, z);
                                uint8_t fl = flags[ci];
                                if (!(fl & F_ACTIVE)) continue;
                                const Real inv = inv_acc(acc_of(fl));
                                const Real G = std::fma(nbr_sum(P, ci, fl), inv, Real(0) - d[ci] * inv);
                                P[ci] = P[ci] + om * (G - P[ci]);
                            }
                }
        }
        // Gradient subtraction on every face whose two cells are fluid.
        std::vector<Real>& U = u[now];
        std::vector<Real>& V = v[now];
        std::vector<Real>& W = w[now];
#pragma omp parallel for schedule(static)
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    long ci = c(x, y, z);
                    uint8_t fl = flags[ci];
                    if (!(fl & F_SELF)) continue;
                    if (x >= 1 && (fl & F_XM)) U[f(x, y, z)] = U[f(x, y, z)] - (P[ci] - P[ci - 1]);
                    if (y >= 1 && (fl & F_YM)) V[f(x, y, z)] = V[f(x, y, z)] - (P[ci] - P[ci - nx]);
                    if (z >= 1 && (fl & F_ZM))
                        W[f(x, y, z)] = W[f(x, y, z)] - (P[ci] - P[ci - (long)nx * ny]);
                }
    }

    // avgU/avgV/avgW (smokeSimulation.cu:409-447): 8-point means, summed in the
    // reference's order.
    Real avgU(const std::vector<Real>& a, int x, int y, int z) const
    {
        Real t = a[f(x, y, z - 1)] + a[f(x + 1, y, z - 1)] + a[f(x, y - 1, z - 1)] +
                 a[f(x + 1, y - 1, z - 1)] + a[f(x, y, z)] + a[f(x + 1, y, z)] + a[f(x, y - 1, z)] +
                 a[f(x + 1, y - 1, z)];
        return t / Real(8);
    }
    Real avgV(const std::vector<Real>& a, int x, int y, int z) const
    {
        Real t = a[f(x, y, z - 1)] + a[f(x - 1, y, z - 1)] + a[f(x, y + 1, z - 1)] +
                 a[f(x - 1, y + 1, z - 1)] + a[f(x, y, z)] + a[f(x - 1, y, z)] + a[f(x, y + 1, z)] +
                 a[f(x - 1, y + 1, z)];
        return t / Real(8);
    }
    Real avgW(const std::vector<Real>& a, int x, int y, int z) const
    {
        Real t = a[f(x, y, z)] + a[f(x - 1, y, z)] + a[f(x, y - 1, z)] + a[f(x - 1, y - 1, z)] +
                 a[f(x, y, z - 1)] + a[f(x - 1, y, z - 1)] + a[f(x, y - 1, z - 1)] +
                 a[f(x - 1, y - 1, z - 1)];
        return t / Real(8);
    }

    // sampleSmoke (smokeSimulation.cu:451-484): positions are clamped with the
    // CELL dims whatever the field; the field's own strides address it.
    Real sample(const std::vector<Real>& a, long gx, long gxy, Real px, Real py, Real pz, Real dx,
                Real dy, Real dz) const
    {
        Real x = std::fmax(std::fmin(px, (Real)(nx - 1)), Real(1));
        Real y = std::fmax(std::fmin(py, (Real)(ny - 1)), Real(1));
        Real z = std::fmax(std::fmin(pz, (Real)(nz - 1)), Real(1));
        int x0 = (int)std::fmin(std::floor(x - dx), (Real)(nx - 1));
        Real xw1 = (x - dx) - (Real)x0;
        Real xw0 = Real(1) - xw1;
        int x1 = (int)std::fmin((Real)(x0 + 1), (Real)(nx - 1));
        int y0 = (int)std::fmin(std::floor(y - dy), (Real)(ny - 1));
        Real yw1 = (y - dy) - (Real)y0;
        Real yw0 = Real(1) - yw1;
        int y1 = (int)std::fmin((Real)(y0 + 1), (Real)(ny - 1));
        int z0 = (int)std::fmin(std::floor(z - dz), (Real)(nz - 1));
        Real zw1 = (z - dz) - (Real)z0;
        Real zw0 = Real(1) - zw1;
        int z1 = (int)std::fmin((Real)(z0 + 1), (Real)(nz - 1));
        auto F = [&](int i, int j, int k) { return a[i + j * gx + k * gxy]; };
        return xw0 * yw0 * zw0 * F(x0, y0, z0) + xw1 * yw0 * zw0 * F(x1, y0, z0) +
               xw0 * yw1 * zw0 * F(x0, y1, z0) + xw1 * yw1 * zw0 * F(x1, y1, z0) +
               xw0 * yw0 * zw1 * F(x0, y0, z1) + xw1 * yw0 * zw1 * F(x1, y0, z1) +
               xw0 * yw1 * zw1 * F(x0, y1, z1) + xw1 * yw1 * zw1 * F(x1, y1, z1);
    }

    // velocityAdvectionU/V/W (smokeSimulation.cu:527-615): read [now], write
    // [past]; faces whose two cells are not both fluid keep stale values.
    void stage_advect_velocity(float dt_f)
    {
        const Real dt = (Real)dt_f;
        const int past = 1 - now;
        const std::vector<Real>& U0 = u[now];
        const std::vector<Real>& V0 = v[now];
        const std::vector<Real>& W0 = w[now];
        std::vector<Real>& U1 = u[past];
        std::vector<Real>& V1 = v[past];
        std::vector<Real>& W1 = w[past];
        const long gxy = Sx * Sy;
#pragma omp parallel for schedule(static)
        for (int z = 1; z < nz; ++z)
            for (int y = 1; y < ny; ++y)
                for (int x = 1; x < nx; ++x) {
                    long ci = c(x, y, z);
                    // U: x in [1,nx-1], y,z in [1,n-2]
                    if (y < ny - 1 && z < nz - 1 && s[ci] != 0 && s[ci - 1] != 0) {
                        Real xt = (Real)x, yt = (Real)(y + 0.5), zt = (Real)(z + 0.5);
                        Real uu = U0[f(x, y, z)];
                        Real vv = avgV(V0, x, y, z);
                        Real ww = avgW(W0, x, y, z);
                        xt = xt - dt * uu;
                        yt = yt - dt * vv;
                        zt = zt - dt * ww;
                        U1[f(x, y, z)] = sample(U0, Sx, gxy, xt, yt, zt, Real(0), Real(0.5), Real(0.5));
                    }
                    // V: y in [1,ny-1], x,z in [1,n-2]
                    if (x < nx - 1 && z < nz - 1 && s[ci] != 0 && s[ci - nx] != 0) {
                        Real xt = (Real)(x + 0.5), yt = (Real)y, zt = (Real)(z + 0.5);
                        Real uu = avgU(U0, x, y, z);
                        Real vv = V0[f(x, y, z)];
                        Real ww = avgW(W0, x, y, z);
                        xt = xt - dt * uu;
                        yt = yt - dt * vv;
                        zt = zt - dt * ww;
                        V1[f(x, y, z)] = sample(V0, Sx, gxy, xt, yt, zt, Real(0.5), Real(0), Real(0.5));
                    }
                    // W: z in [1,nz-1], x,y in [1,n-2]
                    if (x < nx - 1 && y < ny - 1 && s[ci] != 0 && s[ci - (long)nx * ny] != 0) {
                        Real xt = (Real)(x + 0.5), yt = (Real)(y + 0.5), zt = (Real)z;
                        Real uu = avgU(U0, x, y, z);
                        Real vv = avgV(V0, x, y, z);
                        Real ww = W0[f(x, y, z)];
                        xt = xt - dt * uu;
                        yt = yt - dt * vv;
                        zt = zt - dt * ww;
                        W1[f(x, y, z)] = sample(W0, Sx, gxy, xt, yt, zt, Real(0.5), Real(0.5), Real(0));
                    }
                }
    }

    // advectSmoke (smokeSimulation.cu:617-638): the cell velocity is the
    // two-face mean of the POST-advection field [past]; the back-trace is
    // formed in double and rounded.
    void stage_advect_density(float dt_f)
    {
        const Real dt = (Real)dt_f;
        const int past = 1 - now;
        const std::vector<Real>& U = u[past];
        const std::vector<Real>& V = v[past];
        const std::vector<Real>& W = w[past];
        const std::vector<Real>& R0 = rho[now];
        std::vector<Real>& R1 = rho[past];
        const long gxy = (long)nx * ny;
#pragma omp parallel for schedule(static)
        for (int z = 1; z < nz - 1; ++z)
            for (int y = 1; y < ny - 1; ++y)
                for (int x = 1; x < nx - 1; ++x) {
                    long ci = c(x, y, z);
                    if (s[ci] == 0) continue;
                    Real ut = (U[f(x, y, z)] + U[f(x + 1, y, z)]) / Real(2);
                    Real vt = (V[f(x, y, z)] + V[f(x, y + 1, z)]) / Real(2);
                    Real wt = (W[f(x, y, z)] + W[f(x, y, z + 1)]) / Real(2);
                    Real xt = (Real)((double)x + 0.5 - (double)(ut * dt));
                    Real yt = (Real)((double)y + 0.5 - (double)(vt * dt));
                    Real zt = (Real)((double)z + 0.5 - (double)(wt * dt));
                    R1[ci] = sample(R0, nx, gxy, xt, yt, zt, Real(0.5), Real(0.5), Real(0.5));
                }
    }

    void run_stage(int stage, float dt) override
    {
        switch (stage) {
        case STAGE_SCENE: stage_scene(); break;
        case STAGE_FORCE: stage_force(dt); break;
        case STAGE_PROJECT:
            if (prm.projection == PROJ_REFERENCE) project_reference();
            else project_pressure();
            break;
        case STAGE_ADVECT_VEL: stage_advect_velocity(dt); break;
        case STAGE_ADVECT_DENS: stage_advect_density(dt); break;
        default: break;
        }
    }
};

}  // namespace

// ------------------------------------------------------------------ C ABI
// Flat entry points for the test suite (ctypes).  Arrays are exchanged as
// double in the reference layouts described at the top of this file.
extern "C" {

struct orc_params {
    float gravity, buoyancy, density_gate, max_velocity;
    double sor_omega, jacobi_omega;
    int projection, iterations, obstacle_mode, warm_start;
    float vorticity;
    int obstacle_velocity;
};

void* orc_create(int real_bits, int nx, int ny, int nz)
{
    if (nx < 3 || ny < 3 || nz < 3) return nullptr;
    if (real_bits == 64) return static_cast<SimBase*>(new Sim<double>(nx, ny, nz));
    if (real_bits == 32) return static_cast<SimBase*>(new Sim<float>(nx, ny, nz));
    return nullptr;
}

void orc_destroy(void* h) { delete static_cast<SimBase*>(h); }

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_params(void* h, const orc_params* p)
{
    SimBase* s = static_cast<SimBase*>(h);
    s->prm.gravity = p->gravity;
    s->prm.buoyancy = p->buoyancy;
    s->prm.density_gate = p->density_gate;
    s->prm.max_velocity = p->max_velocity;
    s->prm.sor_omega = p->sor_omega;
    s->prm.jacobi_omega = p->jacobi_omega;
    s->prm.projection = p->projection;
    s->prm.iterations = p->iterations;
    s->prm.obstacle_mode = p->obstacle_mode;
    s->prm.warm_start = p->warm_start;
    s->prm.vorticity = p->vorticity;
    s->prm.obstacle_velocity = p->obstacle_velocity;
}

// addObstacle / addSmokeSource (smokeSimulation.cu:88-104): ids are shared
// across both types and index the object list.
int orc_add_object(void* h, int type, float x, float y, float z, float vx, float vy, float vz, float r)
{
    SimBase* s = static_cast<SimBase*>(h);
    Sphere o;
    o.id = (int)s->objects.size();
    o.pos[0] = x; o.pos[1] = y; o.pos[2] = z;
    o.vel[0] = vx; o.vel[1] = vy; o.vel[2] = vz;
    o.r = r;
    o.type = type;
    s->objects.push_back(o);
    return o.id;
}

// updateObjectPos (smokeSimulation.cu:106-109), with a bounds check.
int orc_move_object(void* h, int id, float x, float y, float z)
{
    SimBase* s = static_cast<SimBase*>(h);
    if (id < 0 || id >= (int)s->objects.size()) return -1;
    s->objects[id].pos[0] = x;
    s->objects[id].pos[1] = y;
    s->objects[id].pos[2] = z;
    return 0;
}

// Sets the stored velocity of an object (used by the moving-obstacle
// boundary condition).
int orc_set_object_velocity(void* h, int id, float vx, float vy, float vz)
{
    SimBase* s = static_cast<SimBase*>(h);
    if (id < 0 || id >= (int)s->objects.size()) return -1;
    s->objects[id].vel[0] = vx;
    s->objects[id].vel[1] = vy;
    s->objects[id].vel[2] = vz;
    return 0;
}

// Base mask from a solid array (nonzero = solid).  The mask is rebuilt at
// once, as the next step would.
void orc_set_solids(void* h, const uint8_t* solid)
{
    SimBase* s = static_cast<SimBase*>(h);
    for (long i = 0; i < s->N; ++i) s->base[i] = solid[i] ? 0 : 1;
    s->rebuild_mask();
}

void orc_rebuild_mask(void* h) { static_cast<SimBase*>(h)->rebuild_mask(); }

void orc_set_density(void* h, int half, const double* in) { static_cast<SimBase*>(h)->set_density(half, in); }
void orc_get_density(void* h, int half, double* out) { static_cast<SimBase*>(h)->get_density(half, out); }
void orc_set_velocity(void* h, int comp, int half, const double* in)
{
    static_cast<SimBase*>(h)->set_velocity(comp, half, in);
}
void orc_get_velocity(void* h, int comp, int half, double* out)
{
    static_cast<SimBase*>(h)->get_velocity(comp, half, out);
}
void orc_get_pressure(void* h, double* out) { static_cast<SimBase*>(h)->get_pressure(out); }
void orc_get_rhs(void* h, double* out) { static_cast<SimBase*>(h)->get_rhs(out); }
void orc_get_mask(void* h, uint8_t* out)
{
    SimBase* s = static_cast<SimBase*>(h);
    std::memcpy(out, s->s.data(), s->N);
}
void orc_get_flags(void* h, uint8_t* out)
{
    SimBase* s = static_cast<SimBase*>(h);
    std::memcpy(out, s->flags.data(), s->N);
}
int orc_index_now(void* h) { return static_cast<SimBase*>(h)->now; }

void orc_stage(void* h, int stage, float dt) { static_cast<SimBase*>(h)->run_stage(stage, dt); }

// simulate() (smokeSimulation.cu:774-819) minus the device-to-host copy: the
// output density is half [1 - index_now] afterwards.
void orc_step(void* h, float dt)
{
    SimBase* s = static_cast<SimBase*>(h);
    for (int st = STAGE_SCENE; st <= STAGE_ADVECT_DENS; ++st) s->run_stage(st, dt);
}

void orc_steps(void* h, float dt, int n)
{
    for (int i = 0; i < n; ++i) orc_step(h, dt);
}

}  // extern "C"
