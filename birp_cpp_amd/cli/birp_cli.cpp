// birp_cli.cpp -- command-line renderer over the C++ API (include/*.h), taking the reference CLI's
// flags (reference src/main.cpp:80, getopt string "i:o:x:y:z:a:tcs:h:w:v:b:d:p:q:r:f:e:g:u:j:k:l:")
// with the same defaults and outputs:
//   -i cube.dat   one cube (nearest sampling unless -t)  -> <out><cube.dat>.fits
//   -i dir/       every file of the directory, loaded in parallel and rendered together by one
//                 batched exact launch (Camera::RenderAll); each -> <out><file>.fits
//   -c            the CMEM model -> <out>cmembirp.fits; its Init takes the parameter form whenever
//                 a p list is present, which the default p = {1,1,3,4} always is (main.cpp:276-297)
//   -o out        output prefix (a directory ending in '/')
//   -x -y -z -a   camera position and aim x; -s pixel size (deg); -h -w FOV height and width (deg)
//   -v -b -d -p -q -r -f -e -g -u -j -k -l   CMEM parameters v, b, dipole, p0..p3, B, alpha, beta,
//                 bs, A1, A2, ay_bs, az_bs, density
// Images are bit-identical to the reference's Camera::Render of the same inputs.
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <filesystem>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "Camera.h"
#include "Cmem.h"
#include "DataCube.h"
#include "Helper.h"

namespace {

struct Options {
    std::string input, output;
    bool cmem = false, trilinear = false, batch = false;
    float pos[3] = {5.91813f, 6.74116f, 17.7067f};  // main.cpp:37-40
    float aim = 7.81768f;
    float pixel_deg = 0.25f, fov_h = 36.f, fov_w = 36.f;  // main.cpp:57-63
    std::vector<int> v, b;
    std::vector<float> p = {1.f, 1.f, 3.f, 4.f};  // main.cpp:69
    float dipole = 0, B = 0, alpha = 0, beta = 0, bs = 0, A1 = 0, A2 = 0, ay_bs = 0, az_bs = 0, density = 0;
};

int usage(const char* msg) {
    std::cerr << msg << "\n"
              << "usage: birp -i <cube.dat | dir/> | -c  [-o out/] [-t] [-x X -y Y -z Z -a AIM] [-s deg -h deg -w deg]\n"
              << "            CMEM: [-v v1,v2,v3] [-b b1,b2,b3] [-d dipole] [-p p0,p1,p2,p3] [-q B] [-r alpha] [-f beta]\n"
              << "                  [-e bs] [-g A1] [-u A2] [-j ay_bs] [-k az_bs] [-l density]\n";
    return 3;
}

bool parse(int argc, char** argv, Options& o, std::string& err) {
    int c;
    bool have_input = false;
    while ((c = getopt(argc, argv, "i:o:x:y:z:a:tcs:h:w:v:b:d:p:q:r:f:e:g:u:j:k:l:")) != -1) {
        const std::string arg = optarg ? optarg : "";
        switch (c) {
            case 'i':
                o.input = arg;
                have_input = true;
                if (arg.find(".dat") != std::string::npos) o.batch = false;  // the reference's test (main.cpp:87-97)
                else if (Helper::ends_with(arg, "/")) o.batch = true;
                else {
                    err = "Please provide valid .dat datacube or use -c to use CMEM.";
                    return false;
                }
                break;
            case 'c': o.cmem = true; o.batch = false; break;
            case 'o': o.output = arg; break;
            case 'x': o.pos[0] = std::stof(arg); break;
            case 'y': o.pos[1] = std::stof(arg); break;
            case 'z': o.pos[2] = std::stof(arg); break;
            case 'a': o.aim = std::stof(arg); break;
            case 't': o.trilinear = true; break;
            case 's': o.pixel_deg = std::stof(arg); break;
            case 'h': o.fov_h = std::stof(arg); break;
            case 'w': o.fov_w = std::stof(arg); break;
            case 'v': o.v = Helper::explode_int(arg, ','); break;
            case 'b': o.b = Helper::explode_int(arg, ','); break;
            case 'd': o.dipole = std::stof(arg); break;
            case 'p': o.p = Helper::explode_float(arg, ','); break;
            case 'q': o.B = std::stof(arg); break;
            case 'r': o.alpha = std::stof(arg); break;
            case 'f': o.beta = std::stof(arg); break;
            case 'e': o.bs = std::stof(arg); break;
            case 'g': o.A1 = std::stof(arg); break;
            case 'u': o.A2 = std::stof(arg); break;
            case 'j': o.ay_bs = std::stof(arg); break;
            case 'k': o.az_bs = std::stof(arg); break;
            case 'l': o.density = std::stof(arg); break;
            default: err = "unknown option"; return false;
        }
    }
    if (!have_input && !o.cmem) {
        err = "Please provide valid .dat datacube or use -c to use CMEM.";
        return false;
    }
    return true;
}

std::string base_name(const std::string& path) { return path.substr(path.find_last_of("/\\") + 1); }

void setup(Camera& cam, const Options& o) {
    float x = o.pos[0], y = o.pos[1], z = o.pos[2];
    cam.SetPosition(x, y, z);
    cam.SetAim(o.aim, 0.f, 0.f);
}

// Every file of the directory (in name order), loaded on up to 8 threads, all cameras rendered by
// one Camera::RenderAll.
void run_batch(const Options& o) {
    std::vector<std::string> paths;
    for (const auto& entry : std::filesystem::directory_iterator(o.input))
        if (entry.is_regular_file()) paths.push_back(entry.path().string());
    std::sort(paths.begin(), paths.end());
    std::cout << "Loading " << paths.size() << " datacubes..." << std::flush;
    std::vector<std::unique_ptr<DataCube>> cubes(paths.size());
    std::vector<std::exception_ptr> errs(paths.size());
    std::atomic<size_t> next{0};
    auto worker = [&] {
        for (size_t k; (k = next++) < paths.size();) {
            try {
                cubes[k] = std::make_unique<DataCube>();
                cubes[k]->Load(paths[k], false);
                cubes[k]->SetTrilinear(o.trilinear);
            } catch (...) {
                errs[k] = std::current_exception();
            }
        }
    };
    std::vector<std::thread> pool;
    const size_t nt = std::min<size_t>(8, std::max<size_t>(1, paths.size()));
    for (size_t t = 0; t < nt; t++) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
    for (auto& e : errs)
        if (e) std::rethrow_exception(e);
    std::cout << " loaded.\nRendering..." << std::flush;
    std::vector<Camera> cams;
    cams.reserve(cubes.size());
    std::vector<Camera*> ptrs;
    for (auto& c : cubes) {
        cams.emplace_back(*c, o.pixel_deg, o.fov_h, o.fov_w);
        setup(cams.back(), o);
        ptrs.push_back(&cams.back());
    }
    Camera::RenderAll(ptrs);
    std::cout << " completed.\nExporting..." << std::flush;
    for (size_t k = 0; k < cams.size(); k++) {
        cams[k].ToFITS(o.output + base_name(paths[k]));
        std::cout << "\nExported to " << o.output + base_name(paths[k]) << ".fits" << std::flush;
    }
    std::cout << "\n";
}

void run_single(const Options& o) {
    std::unique_ptr<Space> space;
    std::string base;
    if (!o.cmem) {
        base = base_name(o.input);
        std::cout << "Loading datacube...\t" << std::flush;
        auto cube = std::make_unique<DataCube>();
        std::string path = o.input;
        cube->Init(path, false);
        cube->SetTrilinear(o.trilinear);
        space = std::move(cube);
    } else {
        base = "cmembirp";
        std::cout << "Initialising CMEM...\t" << std::flush;
        auto m = std::make_unique<CMEM>();
        const bool params = o.dipole != 0 || o.B != 0 || o.alpha != 0 || o.beta != 0 || !o.v.empty() || !o.b.empty() ||
                            !o.p.empty() || o.density != 0;
        if (params) {
            if (o.p.size() < 4) throw std::invalid_argument("-p needs four values p0,p1,p2,p3");
            m->Init(true, o.v, o.b, o.dipole, o.p[0], (int)o.p[1], (int)o.p[2], (int)o.p[3], o.B, o.alpha, o.beta, o.bs, o.A1,
                    o.A2, o.ay_bs, o.az_bs, o.density);
        } else {
            m->Init();
        }
        space = std::move(m);
    }
    std::cout << "Rendering...\t\t" << std::flush;
    Camera cam(*space, o.pixel_deg, o.fov_h, o.fov_w);
    setup(cam, o);
    cam.Render();
    std::cout << "Completed rendering.\nExporting...\t\t" << std::flush;
    cam.ToFITS(o.output + base);
    std::cout << "Exported to " << o.output << base << ".fits\n";
}

}  // namespace

int main(int argc, char** argv) {
    const auto t0 = std::chrono::steady_clock::now();
    Options o;
    std::string err;
    if (!parse(argc, argv, o, err)) return usage(err.c_str());
    try {
        if (o.batch) run_batch(o);
        else run_single(o);
    } catch (const std::exception& e) {
        std::cerr << "birp: " << e.what() << "\n";
        return 1;
    }
    const auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "Entire operation took " << ms << "ms or " << ms / 1000 << "s" << std::endl;
    return 0;
}
