// api_tests.cpp -- drives the C++ API (include/DataCube.h, Cmem.h, Camera.h, Helper.h; built into
// libbirp_api.so) for tests/test_api_cpp.py. Each subcommand reads its inputs from argv/stdin and
// prints results (floats as IEEE bit patterns) for the Python side to compare with the golden
// vectors and the oracle. Host-only subcommands: inverse, pointing, cmem_state, load, write_outputs,
// save_load. GPU subcommands: sample, render, cmem_sample, cmem_render, render_twice,
// space_render, render_all.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "Camera.h"
#include "Cmem.h"
#include "DataCube.h"
#include "Helper.h"
#include "birp.h"

static uint32_t bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
static float from_bits(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// A Space that is neither DataCube nor CMEM, like the reference test's CubeFake
// (tests/Camera_test.cpp:25-29).
class CubeFake : public Space {
public:
    void Init() override {}
    float GetSample(float, float, float) override { return 1.f; }
};

// A user Space with negative values: 1 for x > 0.5, -1 elsewhere (rays stop where they cross
// into x <= 0.5); counts its GetSample calls.
class HalfSpace : public Space {
public:
    long long calls = 0;
    void Init() override {}
    float GetSample(float x, float, float) override {
        calls++;
        return x > 0.5f ? 1.f : -1.f;
    }
};

static void print_image(const Camera& cam) {
    std::printf("%d %d\n", cam.Width(), cam.Height());
    for (const auto& row : cam.image)
        for (float v : row) std::printf("%u\n", bits(v));
}

static void init_cmem(CMEM& m, const std::string& which) {
    if (which == "default") {
        m.Init();
    } else if (which == "test_params") {
        // tests/CMEM_test.cpp:51-63, arguments bound exactly as that call binds them
        std::vector<int> v = {400, 0, 0}, b = {0, 0, 4};
        float dipole = 0.f;
        int p1 = 1, p2 = 4, p3 = 6;
        float B = 3.f, alpha = 3.6f, beta = -1.0f;
        m.Init(true, v, b, dipole, p1, p2, p3, B, alpha, beta);
    } else if (which == "cli_params") {
        // main.cpp:285-294 with its defaults: p = {1,1,3,4}, the NULL floats 0
        std::vector<float> p = {1.f, 1.f, 3.f, 4.f};
        m.Init(true, {}, {}, 0.f, p[0], p[1], p[2], p[3], 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f);
    } else {
        throw std::invalid_argument("unknown cmem set " + which);
    }
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string cmd = argv[1];
    try {
        if (cmd == "inverse") {  // 9 uint32 per line -> inverse bits or "throw"
            uint32_t u[9];
            while (std::scanf("%u %u %u %u %u %u %u %u %u", &u[0], &u[1], &u[2], &u[3], &u[4], &u[5], &u[6], &u[7], &u[8]) == 9) {
                std::vector<std::vector<float>> m(3, std::vector<float>(3));
                for (int k = 0; k < 9; k++) m[k / 3][k % 3] = from_bits(u[k]);
                try {
                    auto inv = Helper::GetInverse(m);
                    for (int k = 0; k < 9; k++) std::printf("%u%c", bits(inv[k / 3][k % 3]), k == 8 ? '\n' : ' ');
                } catch (const std::runtime_error& e) {
                    std::printf("throw %s\n", e.what());
                }
            }
            return 0;
        }
        if (cmd == "pointing") {
            // One case per stdin line, floats as bits; one output line per case:
            //   gei                                   -> 16 matrix entries
            //   gsi <ra> <dec>                        -> RaDecToGSI (3)
            //   gse <ra> <dec> <m0..m15>              -> RaDecGSEConversion (3)
            //   twomin <n> <a0..> <nm> <k0..>         -> TwoMin (2)
            //   sky <n> <skysep> <xsky..> <ysky..> <sky1..> <sky2..>   (sets the grid, no output)
            //   xyradec <x> <y>                       -> XYToRaDec on the grid (2)
            //   interp <n> <m> <x..> <y..> <xn..>     -> Interp1 (m), or "throw"
            //   nearest <n> <value> <x..>             -> findNearestNeighbourIndex (int)
            //   sph <x> <y> <z>                       -> XYZToSpherical (3) then XYZToSphericalAlt (3)
            auto rf = [] {
                uint32_t u;
                if (!(std::cin >> u)) throw std::runtime_error("pointing: truncated input");
                return from_bits(u);
            };
            auto rfv = [&](int n) {
                std::vector<float> v(n);
                for (auto& x : v) x = rf();
                return v;
            };
            auto put = [](const std::vector<float>& v) {
                for (size_t k = 0; k < v.size(); k++) std::printf("%u%c", bits(v[k]), k + 1 == v.size() ? '\n' : ' ');
            };
            std::vector<float> xsky, ysky, sky1, sky2;
            float skysep = 0.f;
            std::string op;
            while (std::cin >> op) {
                if (op == "gei") {
                    std::vector<float> flat;
                    for (const auto& r : Helper::Gei2GseTransforms()) flat.insert(flat.end(), r.begin(), r.end());
                    put(flat);
                } else if (op == "gsi") {
                    float ra = rf(), dec = rf();
                    put(Helper::RaDecToGSI(ra, dec));
                } else if (op == "gse") {
                    float ra = rf(), dec = rf();
                    std::vector<std::vector<float>> m(4);
                    for (auto& r : m) r = rfv(4);
                    put(Helper::RaDecGSEConversion({ra, dec}, m));
                } else if (op == "twomin") {
                    int n;
                    std::cin >> n;
                    auto a = rfv(n);
                    int nm;
                    std::cin >> nm;
                    std::vector<int> mask(nm);
                    for (auto& k : mask) std::cin >> k;
                    put(Helper::TwoMin(a, mask));
                } else if (op == "sky") {
                    int n;
                    std::cin >> n;
                    skysep = rf();
                    xsky = rfv(n);
                    ysky = rfv(n);
                    sky1 = rfv(n);
                    sky2 = rfv(n);
                } else if (op == "xyradec") {
                    float x = rf(), y = rf();
                    put(Helper::XYToRaDec(x, y, xsky, ysky, sky1, sky2, skysep));
                } else if (op == "interp") {
                    int n, m;
                    std::cin >> n >> m;
                    auto x = rfv(n), y = rfv(n), xn = rfv(m);
                    try {
                        put(Helper::Interp1(x, y, xn));
                    } catch (const std::out_of_range&) {
                        std::printf("throw\n");
                    }
                } else if (op == "nearest") {
                    int n;
                    std::cin >> n;
                    float v = rf();
                    auto x = rfv(n);
                    std::printf("%d\n", Helper::findNearestNeighbourIndex(v, x));
                } else if (op == "sph") {
                    float p[3] = {rf(), rf(), rf()};
                    auto a = Helper::XYZToSpherical(p), b = Helper::XYZToSphericalAlt(p);
                    a.insert(a.end(), b.begin(), b.end());
                    put(a);
                } else {
                    throw std::invalid_argument("pointing: unknown op " + op);
                }
            }
            return 0;
        }
        if (cmd == "cmem_state") {  // post-Init members as bits
            CMEM m;
            init_cmem(m, argv[2]);
            birp_cmem_state s;
            m.State(&s);
            const float* f = &s.p0;
            for (size_t k = 0; k < sizeof(s) / sizeof(float); k++) std::printf("%u\n", bits(f[k]));
            return 0;
        }
        if (cmd == "load") {  // DataCube::Load: size, spacing, origin, then the values
            DataCube c;
            std::string path = argv[2];
            c.Init(path, false);
            std::printf("%d %d %d\n", c.size.x, c.size.y, c.size.z);
            std::printf("%u %u %u\n", bits(c.spacing[0]), bits(c.spacing[1]), bits(c.spacing[2]));
            std::printf("%u %u %u\n", bits(c.origin_coords[0]), bits(c.origin_coords[1]), bits(c.origin_coords[2]));
            for (const auto& s : c.slices)
                for (const auto& r : s)
                    for (float v : r) std::printf("%u\n", bits(v));
            return 0;
        }
        if (cmd == "save_load") {  // Load, Save, Load again: identical cube
            DataCube a, b;
            std::string path = argv[2], out = argv[3];
            a.Init(path, false);
            a.Save(out);
            b.Init(out, false);
            bool same = a.coords_x == b.coords_x && a.coords_y == b.coords_y && a.coords_z == b.coords_z && a.slices == b.slices;
            std::printf("%s\n", same ? "same" : "different");
            return same ? 0 : 1;
        }
        if (cmd == "bad_load") {  // a missing or truncated file throws std::invalid_argument
            DataCube c;
            std::string path = argv[2];
            try {
                c.Init(path, false);
                std::printf("no throw\n");
                return 1;
            } catch (const std::invalid_argument& e) {
                std::printf("invalid_argument %s\n", e.what());
                return 0;
            }
        }
        if (cmd == "write_outputs") {  // ToDat / ToFITS of a known image (no render)
            CubeFake fake;
            Camera cam(fake, 0.25f, 2.0f, 3.0f);  // 12 x 8
            float x = 1.5f, y = -2.f, z = 3.25f;
            cam.SetPosition(x, y, z);
            cam.SetAim(7.f, 0.f, 0.f);
            for (int j = 0; j < cam.Height(); j++)
                for (int i = 0; i < cam.Width(); i++) cam.image[j][i] = 0.125f * i - 3.5f * j + 1e-3f * i * j;
            std::string base = argv[2];
            if (cam.ToDat(base) != 1 || cam.ToFITS(base) != 1) return 1;
            print_image(cam);
            return 0;
        }
        // ------------------------------------------------------------------ GPU subcommands
        if (cmd == "sample") {  // DataCube::GetSample at points read as "x y z" bits
            DataCube c;
            std::string path = argv[2];
            c.Init(path, false);
            c.SetTrilinear(std::atoi(argv[3]) != 0);
            uint32_t u[3];
            while (std::scanf("%u %u %u", &u[0], &u[1], &u[2]) == 3)
                std::printf("%u\n", bits(c.GetSample(from_bits(u[0]), from_bits(u[1]), from_bits(u[2]))));
            return 0;
        }
        if (cmd == "render" || cmd == "render_twice") {
            // render <dat> <trilinear> <pixel|W> <H or 0> <fov_h> <fov_w> <px> <py> <pz> <ax> <L> <n> <semantics>
            DataCube c;
            std::string path = argv[2];
            c.Init(path, false);
            c.SetTrilinear(std::atoi(argv[3]) != 0);
            int H = std::atoi(argv[5]);
            float fov_h = std::atof(argv[6]), fov_w = std::atof(argv[7]);
            Camera cam = H > 0 ? Camera(c, std::atoi(argv[4]), H, fov_h, fov_w) : Camera(c, (float)std::atof(argv[4]), fov_h, fov_w);
            float px = std::atof(argv[8]), py = std::atof(argv[9]), pz = std::atof(argv[10]);
            cam.SetPosition(px, py, pz);
            cam.SetAim(std::atof(argv[11]), 0.f, 0.f);
            cam.SetRay(std::atof(argv[12]), std::atoi(argv[13]));
            cam.SetSemantics(std::atoi(argv[14]));
            cam.Render();
            if (cmd == "render_twice") cam.Render();
            print_image(cam);
            std::fprintf(stderr, "in_cube %lld kernel_ms %.4f\n", cam.LastSamplesInCube(), cam.LastKernelMs());
            return 0;
        }
        if (cmd == "cmem_sample") {
            CMEM m;
            init_cmem(m, argv[2]);
            uint32_t u[3];
            while (std::scanf("%u %u %u", &u[0], &u[1], &u[2]) == 3)
                std::printf("%u\n", bits(m.GetSample(from_bits(u[0]), from_bits(u[1]), from_bits(u[2]))));
            return 0;
        }
        if (cmd == "cmem_render") {  // the reference's commented-out CMEM render test pose (Camera_test.cpp:122-158)
            CMEM m;
            init_cmem(m, argv[2]);
            Camera cam(m, 0.25f, 36.f, 36.f);
            float x = 1.f, y = 1.f, z = 1.f;
            cam.SetPosition(x, y, z);
            cam.SetAim(1, 0, 0);
            cam.Render();
            print_image(cam);
            return 0;
        }
        if (cmd == "space_render") {
            // space_render <fake|halfspace> <outdir or -> <pixel> <fov_h> <fov_w> <px> <py> <pz> <ax> <L> <n> <times>:
            // the reference test's CubeFake scenario (Camera_test.cpp:31-73) and a user Space with
            // negative values (1 for x > x0, -1 elsewhere), rendered <times> times (Render accumulates),
            // then ToDat / ToFITS into <outdir>
            const std::string kind = argv[2], outdir = argv[3];
            CubeFake fake;
            HalfSpace half;
            Space* space = kind == "fake" ? static_cast<Space*>(&fake) : static_cast<Space*>(&half);
            Camera cam(*space, (float)std::atof(argv[4]), (float)std::atof(argv[5]), (float)std::atof(argv[6]));
            float px = std::atof(argv[7]), py = std::atof(argv[8]), pz = std::atof(argv[9]);
            cam.SetPosition(px, py, pz);
            cam.SetAim(std::atof(argv[10]), 0.f, 0.f);
            cam.SetRay(std::atof(argv[11]), std::atoi(argv[12]));
            const int times = std::atoi(argv[13]);
            for (int t = 0; t < times; t++) cam.Render();
            if (outdir != "-") {
                cam.ToDat(outdir + "/testout");
                cam.ToFITS(outdir + "/testout");
            }
            print_image(cam);
            std::fprintf(stderr, "calls %lld\n", half.calls);
            return 0;
        }
        if (cmd == "render_all") {
            // render_all <L> <n> <semantics> <dat>...: one camera per cube (trilinear, sizes and poses
            // varying), then a CMEM camera; Camera::RenderAll twice, then the same cameras again with
            // Render() twice each; prints every image of both, then the in-cube counts
            std::vector<DataCube> cubes(argc - 5);
            CMEM m;
            m.Init();
            for (size_t k = 0; k < cubes.size(); k++) {
                std::string path = argv[5 + k];
                cubes[k].Init(path, false);
                cubes[k].SetTrilinear(true);
            }
            auto make = [&](std::vector<Camera>& cams) {
                cams.reserve(cubes.size() + 1);
                for (size_t k = 0; k < cubes.size(); k++) cams.emplace_back(cubes[k], 48 + 8 * (int)k, 32, 20.f, 30.f);
                cams.emplace_back(m, 40, 24, 20.f, 30.f);
                for (size_t k = 0; k < cams.size(); k++) {
                    float px = 0.f + k, py = -20.f + 3.f * k, pz = 10.f;
                    cams[k].SetPosition(px, py, pz);
                    cams[k].SetAim(10.f, 0.f, 0.f);
                    cams[k].SetRay(std::atof(argv[2]), std::atoi(argv[3]));
                    cams[k].SetSemantics(std::atoi(argv[4]));
                }
            };
            std::vector<Camera> batched, single;
            make(batched);
            make(single);
            std::vector<Camera*> ptrs;
            for (auto& c : batched) ptrs.push_back(&c);
            Camera::RenderAll(ptrs);
            Camera::RenderAll(ptrs);
            for (auto& c : single) {
                c.Render();
                c.Render();
            }
            for (auto* set : {&batched, &single})
                for (auto& c : *set) print_image(c);
            for (auto* set : {&batched, &single})
                for (auto& c : *set) std::printf("%lld\n", c.LastSamplesInCube());
            return 0;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "exception: %s\n", e.what());
        return 3;
    }
    return 2;
}
