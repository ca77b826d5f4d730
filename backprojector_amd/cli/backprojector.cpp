// backprojector: NLOS dataset (HDF5) -> backprojection volume (HDF5), the command line of the
// reference (main.cpp:29-187) over libbpx.  Flags, defaults, dataset field names, bounce
// selection, data-order handling, resolution rescale, squeeze and output format follow the
// reference; the reconstruction runs on the GPU through the C ABI of include/bpx.h.
//
//   backprojector <file.hdf5> [-b N] [-r x,y,z] [-c x,y,z] [-s x,y,z] [-d x,y,z] [-o out.hdf5]
//                 [--cpu] [--octree] [--phasor] [-w wavelength] [--dry-run]
//
// --dry-run (no reference counterpart) reads the dataset and prints the derived
// reconstruction parameters as JSON without touching the GPU.
#include <hdf5.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "bpx.h"

namespace {

struct Array {
    std::vector<hsize_t> dims;
    std::vector<float> f;   // float fields
    std::vector<int> i;     // integer fields
    size_t size() const {
        size_t n = 1;
        for (hsize_t d : dims) n *= (size_t)d;
        return n;
    }
};

[[noreturn]] void die(const std::string& msg) { throw std::runtime_error(msg); }

struct H5Handle {  // closes an HDF5 identifier of any kind
    hid_t id;
    herr_t (*close)(hid_t);
    H5Handle(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~H5Handle() {
        if (id >= 0) close(id);
    }
    H5Handle(const H5Handle&) = delete;
    H5Handle& operator=(const H5Handle&) = delete;
};

std::vector<hsize_t> dims_of(hid_t space) {
    const int rank = H5Sget_simple_extent_ndims(space);
    if (rank < 0) die("cannot read a dataspace");
    std::vector<hsize_t> d((size_t)rank);
    if (rank) H5Sget_simple_extent_dims(space, d.data(), nullptr);
    return d;
}

// Whole field as float or int32 (dataset_loader.hpp:53-76: integer class -> int32, else float).
Array load_field(hid_t file, const char* name) {
    if (H5Lexists(file, name, H5P_DEFAULT) <= 0) die(std::string("dataset field '") + name + "' is missing");
    H5Handle ds(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose);
    if (ds.id < 0) die(std::string("cannot open field '") + name + "'");
    H5Handle space(H5Dget_space(ds.id), H5Sclose);
    H5Handle type(H5Dget_type(ds.id), H5Tclose);
    Array a;
    a.dims = dims_of(space.id);
    if (H5Tget_class(type.id) == H5T_INTEGER) {
        a.i.resize(std::max<size_t>(a.size(), 1));
        if (H5Dread(ds.id, H5T_NATIVE_INT32, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.i.data()) < 0)
            die(std::string("cannot read field '") + name + "'");
        a.f.assign(a.i.begin(), a.i.end());
    } else {
        a.f.resize(std::max<size_t>(a.size(), 1));
        if (H5Dread(ds.id, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.f.data()) < 0)
            die(std::string("cannot read field '") + name + "'");
    }
    return a;
}

std::string string_attribute(hid_t obj, const char* name) {
    if (H5Aexists(obj, name) <= 0) return "";
    H5Handle att(H5Aopen(obj, name, H5P_DEFAULT), H5Aclose);
    H5Handle type(H5Aget_type(att.id), H5Tclose);
    if (H5Tget_class(type.id) != H5T_STRING) return "";
    if (H5Tis_variable_str(type.id) > 0) {
        H5Handle mem(H5Tcopy(H5T_C_S1), H5Tclose);
        H5Tset_size(mem.id, H5T_VARIABLE);
        char* s = nullptr;
        if (H5Aread(att.id, mem.id, &s) < 0 || !s) return "";
        std::string out(s);
        H5free_memory(s);
        return out;
    }
    const size_t n = H5Tget_size(type.id);
    std::vector<char> buf(n + 1, 0);
    H5Handle mem(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(mem.id, n);
    if (H5Aread(att.id, mem.id, buf.data()) < 0) return "";
    return std::string(buf.data());
}

enum Order { kRowMajor, kColumnMajor };

// The data field with the chosen bounces selected on the bounce axis and summed
// (dataset_loader.hpp:87-141: axis rank-3 row-major / 2 column-major, offset
// bounce - (3 - third_bounce), third_bounce attribute default 1).
Array load_transient(hid_t file, const std::vector<uint32_t>& bounces, Order order) {
    if (bounces.empty()) die("no bounces selected");
    if (H5Lexists(file, "data", H5P_DEFAULT) <= 0) die("dataset field 'data' is missing");
    H5Handle ds(H5Dopen2(file, "data", H5P_DEFAULT), H5Dclose);
    int third_bounce = 1;
    if (H5Aexists(ds.id, "third_bounce") > 0) {
        H5Handle att(H5Aopen(ds.id, "third_bounce", H5P_DEFAULT), H5Aclose);
        H5Aread(att.id, H5T_NATIVE_INT, &third_bounce);
    }
    H5Handle space(H5Dget_space(ds.id), H5Sclose);
    std::vector<hsize_t> dims = dims_of(space.id);
    const int rank = (int)dims.size();
    if (rank < 3) die("the data field needs at least 3 axes");
    const size_t axis = order == kRowMajor ? (size_t)(rank - 3) : 2u;
    std::vector<hsize_t> start(dims.size(), 0), count = dims;
    count[axis] = 1;
    H5Sselect_none(space.id);
    for (uint32_t b : bounces) {
        const long off = (long)b - (3 - third_bounce);
        if (off < 0 || (hsize_t)off >= dims[axis])
            die("bounce " + std::to_string(b) + " is not in the dataset");
        start[axis] = (hsize_t)off;
        if (H5Sselect_hyperslab(space.id, H5S_SELECT_OR, start.data(), nullptr, count.data(), nullptr) < 0)
            die("cannot select bounces");
    }
    std::vector<hsize_t> mdims = dims;
    mdims[axis] = bounces.size();
    H5Handle mspace(H5Screate_simple(rank, mdims.data(), nullptr), H5Sclose);
    Array sel;
    sel.dims = mdims;
    sel.f.resize(sel.size());
    if (H5Dread(ds.id, H5T_NATIVE_FLOAT, mspace.id, space.id, H5P_DEFAULT, sel.f.data()) < 0)
        die("cannot read the data field");
    if (bounces.size() == 1) return sel;
    // sum over the bounce axis, keeping it with extent 1
    size_t outer = 1, inner = 1;
    for (size_t k = 0; k < axis; ++k) outer *= (size_t)mdims[k];
    for (size_t k = axis + 1; k < mdims.size(); ++k) inner *= (size_t)mdims[k];
    Array sum;
    sum.dims = mdims;
    sum.dims[axis] = 1;
    sum.f.assign(outer * inner, 0.0f);
    for (size_t o = 0; o < outer; ++o)
        for (size_t b = 0; b < bounces.size(); ++b)
            for (size_t i = 0; i < inner; ++i)
                sum.f[o * inner + i] += sel.f[(o * bounces.size() + b) * inner + i];
    return sum;
}

// Reverses every axis (xt::transpose without permutation, dataset_loader.hpp:222-230).
Array transpose_all(const Array& a) {
    Array t;
    t.dims.assign(a.dims.rbegin(), a.dims.rend());
    const size_t n = a.size(), r = a.dims.size();
    t.f.resize(n);
    std::vector<size_t> idx(r, 0);
    for (size_t lin = 0; lin < n; ++lin) {
        // idx = multi-index of lin in a (row-major); destination index uses reversed axes
        size_t rem = lin;
        for (size_t k = r; k-- > 0;) {
            idx[k] = rem % a.dims[k];
            rem /= a.dims[k];
        }
        size_t dst = 0;
        for (size_t k = r; k-- > 0;) dst = dst * t.dims[r - 1 - k] + idx[k];
        t.f[dst] = a.f[lin];
    }
    return t;
}

bool parse_triple(const std::string& s, double out[3]) {
    // comma-separated, like ArrayReader (main.cpp:10-26)
    size_t pos = 0;
    for (int k = 0; k < 3; ++k) {
        size_t used = 0;
        try {
            out[k] = std::stod(s.substr(pos), &used);
        } catch (...) {
            return false;
        }
        pos += used + 1;
        if (pos > s.size()) pos = 0;
    }
    return true;
}

void save_volume(const std::string& path, const std::vector<float>& vol, const uint32_t n[3]) {
    // dataset "voxelVolume", chunks min(32, extent), deflate 4, NaN fill (octree_volume.hpp:201-223)
    H5Handle file(H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) die("cannot create " + path);
    const hsize_t dims[3] = {n[0], n[1], n[2]};
    const hsize_t chunk[3] = {std::min<hsize_t>(32, n[0]), std::min<hsize_t>(32, n[1]), std::min<hsize_t>(32, n[2])};
    H5Handle plist(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
    const float fill = NAN;
    H5Pset_chunk(plist.id, 3, chunk);
    H5Pset_deflate(plist.id, 4);
    H5Pset_fill_value(plist.id, H5T_NATIVE_FLOAT, &fill);
    H5Handle space(H5Screate_simple(3, dims, nullptr), H5Sclose);
    H5Handle ds(H5Dcreate2(file.id, "voxelVolume", H5T_NATIVE_FLOAT, space.id, H5P_DEFAULT, plist.id, H5P_DEFAULT),
                H5Dclose);
    if (ds.id < 0) die("cannot create voxelVolume in " + path);
    if (H5Dwrite(ds.id, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, vol.data()) < 0)
        die("cannot write " + path);
}

void check(bpx_status s) {
    if (s != BPX_OK) die(std::string("libbpx: ") + bpx_last_error());
}

std::string json_dims(const std::vector<hsize_t>& d) {
    std::string s = "[";
    for (size_t k = 0; k < d.size(); ++k) s += (k ? ", " : "") + std::to_string(d[k]);
    return s + "]";
}

void usage() {
    std::printf(
        "  backprojector {filename} {OPTIONS}\n\n"
        "    NLOS dataset backprojector.\n\n"
        "  OPTIONS:\n"
        "      -h, --help                        Display this help menu\n"
        "      filename                          The input file name.\n"
        "      -b[highest_bounce], --max-bounce=[highest_bounce]  Maximum bounces (from 3, default 3)\n"
        "      -r[x,y,z], --voxel-resolution=[x,y,z]             Resolution of the volume (32,32,32)\n"
        "      -c[x,y,z], --volume-position=[x,y,z]              Centre of the volume (dataset)\n"
        "      -s[x,y,z], --volume-size=[x,y,z]                  Size of the volume (dataset)\n"
        "      -d[x,y,z], --volume-direction=[x,y,z]             Not implemented (ignored)\n"
        "      -o[outfile], --output=[outfile]                   Output file\n"
        "      --cpu                             CPU-path semantics (windowed)\n"
        "      --octree                          Octree refinement\n"
        "      --phasor                          Phasor field reconstruction\n"
        "      -w[wavelength], --wl=[wavelength], --wavelength=[wavelength]  Phasor wavelength (0.005)\n"
        "      --dry-run                         Print the derived parameters and stop\n");
}

struct Options {
    std::string filename, outfile;
    uint32_t highest_bounce = 3;
    double res[3] = {32, 32, 32}, pos[3] = {NAN, NAN, NAN}, size[3] = {NAN, NAN, NAN};
    bool cpu = false, octree = false, phasor = false, dry_run = false;
    float wavelength = 0.005f;
};

Options parse(int argc, const char* argv[]) {
    Options o;
    std::vector<std::string> pos_args;
    auto value = [&](int& i, const std::string& a, const std::string& s, const std::string& l) -> std::string {
        // -xVALUE, -x VALUE, --long=VALUE, --long VALUE
        if (a.rfind(l + "=", 0) == 0) return a.substr(l.size() + 1);
        if (a == s || a == l) {
            if (i + 1 >= argc) die("flag " + a + " needs a value");
            return argv[++i];
        }
        return a.substr(s.size());
    };
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto is = [&](const std::string& s, const std::string& l) {
            return a == s || a == l || a.rfind(l + "=", 0) == 0 || (s.size() == 2 && a.rfind(s, 0) == 0 && a.size() > 2 && a[1] != '-');
        };
        if (a == "-h" || a == "--help") {
            usage();
            std::exit(1);
        } else if (a == "--cpu") {
            o.cpu = true;
        } else if (a == "--octree") {
            o.octree = true;
        } else if (a == "--phasor") {
            o.phasor = true;
        } else if (a == "--dry-run") {
            o.dry_run = true;
        } else if (is("-b", "--max-bounce")) {
            o.highest_bounce = (uint32_t)std::stoul(value(i, a, "-b", "--max-bounce"));
        } else if (is("-r", "--voxel-resolution")) {
            if (!parse_triple(value(i, a, "-r", "--voxel-resolution"), o.res)) die("bad --voxel-resolution");
        } else if (is("-c", "--volume-position")) {
            if (!parse_triple(value(i, a, "-c", "--volume-position"), o.pos)) die("bad --volume-position");
        } else if (is("-s", "--volume-size")) {
            if (!parse_triple(value(i, a, "-s", "--volume-size"), o.size)) die("bad --volume-size");
        } else if (is("-d", "--volume-direction")) {
            double ignored[3];
            (void)parse_triple(value(i, a, "-d", "--volume-direction"), ignored);
        } else if (is("-o", "--output")) {
            o.outfile = value(i, a, "-o", "--output");
        } else if (a.rfind("--wl", 0) == 0 || a.rfind("--wavelength", 0) == 0 || a.rfind("-w", 0) == 0) {
            const std::string l = a.rfind("--wavelength", 0) == 0 ? "--wavelength" : (a.rfind("--wl", 0) == 0 ? "--wl" : "-w");
            o.wavelength = std::stof(value(i, a, l == "-w" ? "-w" : l, l));
        } else if (!a.empty() && a[0] == '-') {
            die("unknown flag " + a);
        } else {
            pos_args.push_back(a);
        }
    }
    if (pos_args.size() != 1) die("expected exactly one input file name");
    o.filename = pos_args[0];
    return o;
}

int run(int argc, const char* argv[]) {
    const Options o = parse(argc, argv);
    std::vector<uint32_t> bounces;
    for (uint32_t b = 3; b <= o.highest_bounce; ++b) bounces.push_back(b);

    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);  // errors become messages, not HDF5 stack dumps
    H5Handle file(H5Fopen(o.filename.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) die("cannot open " + o.filename);
    // data order attribute (dataset_loader.hpp:177-190); NLOSDataset defaults to column-major
    Order file_order = kColumnMajor;
    const std::string ord = string_attribute(file.id, "data order");
    if (ord == "row-major") file_order = kRowMajor;
    Array data = load_transient(file.id, bounces, file_order);
    Array cam_grid = load_field(file.id, "cameraGridPositions");
    Array cam_pos = load_field(file.id, "cameraPosition");
    Array las_grid = load_field(file.id, "laserGridPositions");
    Array las_pos = load_field(file.id, "laserPosition");
    Array vol_pos = load_field(file.id, "hiddenVolumePosition");
    Array vol_size = load_field(file.id, "hiddenVolumeSize");
    Array t0a = load_field(file.id, "t0");
    Array dta = load_field(file.id, "deltaT");
    Array conf = load_field(file.id, "isConfocal");
    if (cam_pos.size() < 3 || las_pos.size() < 3 || vol_pos.size() < 3 || vol_size.size() < 3 || t0a.size() < 1 ||
        dta.size() < 1 || conf.size() < 1)
        die("malformed dataset scalars");
    const bool confocal = conf.f[0] != 0.0f;
    if (file_order != kRowMajor) {  // the CLI forces row-major (main.cpp:118)
        data = transpose_all(data);
        cam_grid = transpose_all(cam_grid);
        las_grid = transpose_all(las_grid);
    }
    const float deltaT = dta.f[0], t0 = t0a.f[0];
    float vpos[3], vsize[3];
    const bool pos_nan = std::isnan(o.pos[0]) || std::isnan(o.pos[1]) || std::isnan(o.pos[2]);
    const bool size_nan = std::isnan(o.size[0]) || std::isnan(o.size[1]) || std::isnan(o.size[2]);
    for (int k = 0; k < 3; ++k) {
        vpos[k] = pos_nan ? vol_pos.f[k] : (float)o.pos[k];
        vsize[k] = size_nan ? vol_size.f[k] : (float)o.size[k];
    }
    // equal resolutions are scaled to the volume's aspect (main.cpp:128-131), truncated to integers
    uint32_t N[3];
    for (int k = 0; k < 3; ++k) N[k] = (uint32_t)o.res[k];
    if (N[0] == N[1] && N[1] == N[2]) {
        const float mx = std::max(vsize[0], std::max(vsize[1], vsize[2]));
        for (int k = 0; k < 3; ++k) N[k] = (uint32_t)((float)N[k] * vsize[k] / mx);
    }
    // squeeze bounce and channel (main.cpp:139-141): confocal [lx][ly][b][T][c] -> [lx][ly][T],
    // exhaustive [lx][ly][cx][cy][b][T][c] -> [lx][ly][cx][cy][T] (first bounce / channel)
    const size_t r = data.dims.size();
    if (r != (confocal ? 5u : 7u))
        die("the data field has " + std::to_string(r) + " axes; expected " + (confocal ? "5" : "7"));
    const size_t T = (size_t)data.dims[r - 2], C = (size_t)data.dims[r - 1], B = (size_t)data.dims[r - 3];
    size_t rows = 1;
    for (size_t k = 0; k + 3 < r; ++k) rows *= (size_t)data.dims[k];
    std::vector<float> transient(rows * T);
    for (size_t row = 0; row < rows; ++row)
        for (size_t t = 0; t < T; ++t) transient[row * T + t] = data.f[(row * B * T + t) * C];
    std::vector<int64_t> tshape;
    for (size_t k = 0; k + 3 < r; ++k) tshape.push_back((int64_t)data.dims[k]);
    tshape.push_back((int64_t)T);
    auto shape64 = [](const Array& a) {
        std::vector<int64_t> s;
        for (hsize_t d : a.dims) s.push_back((int64_t)d);
        return s;
    };
    const std::vector<int64_t> cshape = shape64(cam_grid), lshape = shape64(las_grid);
    std::string outfile = o.outfile;
    if (outfile.empty()) outfile = o.filename.substr(0, o.filename.find(".hdf5")) + "_recon.hdf5";

    if (o.dry_run) {
        std::vector<hsize_t> ts(tshape.begin(), tshape.end());
        std::printf(
            "{\"confocal\": %s, \"data_order\": \"%s\", \"bounces\": %zu, \"transient_shape\": %s, "
            "\"transient_sum\": %.9g, \"camera_grid_shape\": %s, \"laser_grid_shape\": %s, \"t0\": %.9g, "
            "\"deltaT\": %.9g, \"volume_position\": [%.9g, %.9g, %.9g], \"volume_size\": [%.9g, %.9g, %.9g], "
            "\"voxels\": [%u, %u, %u], \"mode\": \"%s\", \"wavelength\": %.9g, \"output\": \"%s\"}\n",
            confocal ? "true" : "false", file_order == kRowMajor ? "row-major" : "column-major", bounces.size(),
            json_dims(ts).c_str(), [&] {
                double s = 0;
                for (float v : transient) s += v;
                return s;
            }(),
            json_dims(cam_grid.dims).c_str(), json_dims(las_grid.dims).c_str(), t0, deltaT, vpos[0], vpos[1], vpos[2],
            vsize[0], vsize[1], vsize[2], N[0], N[1], N[2],
            o.phasor ? "phasor" : (o.octree ? "octree" : (o.cpu ? "cpu" : "gpu")), o.wavelength, outfile.c_str());
        return 0;
    }

    const size_t nvox = (size_t)N[0] * N[1] * N[2];
    std::vector<float> volume(nvox, 0.0f);
    const float zero3[3] = {0, 0, 0};
    if (o.phasor) {
        std::vector<float> cv(2 * nvox);
        check(bpx_nlos_phasor(transient.data(), tshape.data(), (int)tshape.size(), cam_grid.f.data(), cshape.data(),
                              las_grid.f.data(), lshape.data(), cam_pos.f.data(), las_pos.f.data(), t0, deltaT,
                              confocal ? 1 : 0, vpos, vsize, N, o.wavelength, 1, o.cpu ? 1 : 0, cv.data()));
        for (size_t i = 0; i < nvox; ++i) volume[i] = std::abs(std::complex<float>(cv[2 * i], cv[2 * i + 1]));
    } else if (o.octree && !o.cpu) {
        bpx_nlos_prepared prep{};
        check(bpx_nlos_prepare(transient.data(), tshape.data(), (int)tshape.size(), cam_grid.f.data(), cshape.data(),
                               las_grid.f.data(), lshape.data(), cam_pos.f.data(), las_pos.f.data(), t0, deltaT,
                               confocal ? 1 : 0, vpos, vsize, N, zero3, 1, &prep));
        // levels: as many halvings as every side allows (at most 4); 0 -> dense
        uint32_t levels = 0;
        while (levels < 4 && N[0] % (2u << levels) == 0 && N[1] % (2u << levels) == 0 && N[2] % (2u << levels) == 0)
            ++levels;
        bpx_status st;
        if (confocal && levels > 0) {
            bpx_octree_options oo{levels, 0.25f, 0.5f, 0u, {}};
            st = bpx_octree_backprojection(prep.chunk, prep.rows * prep.Tc, prep.Tc, prep.pairs, prep.num_pairs,
                                           volume.data(), N, prep.zero, prep.inc, prep.chunked_t0, deltaT, &oo,
                                           nullptr);
        } else {
            st = bpx_backprojection(prep.chunk, prep.rows * prep.Tc, prep.Tc, prep.pairs, prep.num_pairs, volume.data(),
                                    N, prep.zero, prep.inc, prep.chunked_t0, deltaT);
        }
        bpx_nlos_prepared_free(&prep);
        check(st);
    } else {
        check(bpx_nlos_backproject(transient.data(), tshape.data(), (int)tshape.size(), cam_grid.f.data(),
                                   cshape.data(), las_grid.f.data(), lshape.data(), cam_pos.f.data(), las_pos.f.data(),
                                   t0, deltaT, confocal ? 1 : 0, vpos, vsize, N, zero3, 1, o.cpu ? 1 : 0,
                                   volume.data()));
    }
    save_volume(outfile, volume, N);
    std::printf("Backprojected to %s\n", outfile.c_str());
    return 0;
}

}  // namespace

int main(int argc, const char* argv[]) {
    try {
        return run(argc, argv);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "backprojector: %s\n", e.what());
        return 1;
    }
}
