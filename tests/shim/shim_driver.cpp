// shim_driver.cpp -- test program for the reference-named shim (endtoend_amd/shim): it fills the
// reference's types the way dosecalc-beam/main.cu and dosecalc-beamlet/main.cu do and calls the
// reference's entry points in their order.  TEST INFRASTRUCTURE.
//
//   shim_driver <case dir> <output>
// The case directory (written by tests/test_gpu_shim.py) holds case.txt (scalars and one "beam"
// line per beam) and raw little-endian float32 files: density, kernel, radii, theta, phi, spectrum
// (fluence, energy, mu_en, mu rows) and fmap_<i> per beam.  Built without CCCS_SHIM_BEAMLET it
// writes the summed dose as raw float32 to <output> (dosecalc-beam/main.cu:313-381); with it, the
// coefficient file <output> (dosecalc-beamlet/main.cu:392-556).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "reference_shim.hpp"

#ifdef CCCS_SHIM_BEAMLET
SparsifyManager sparsifymanager;  // dosecalc-beamlet/main.cu:49
#endif

namespace {
std::vector<float> read_raw(const std::string& path, size_t n) {
    std::vector<float> v(n);
    std::ifstream f(path, std::ios::binary);
    if (!f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(float))))
        throw std::runtime_error("cannot read " + path);
    return v;
}
}  // namespace

int main(int argc, char** argv) {
    if (argc != 3 && argc != 4) {
        std::fprintf(stderr, "usage: %s <case dir> <output> [central|per_beamlet]\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1], out = argv[2];
    const bool per_beamlet = argc == 4 && std::string(argv[3]) == "per_beamlet";
    (void)per_beamlet;  // the beamlet program only
    std::ifstream cf(dir + "/case.txt");
    std::map<std::string, std::vector<double>> kv;
    std::vector<std::vector<double>> beam_lines;
    for (std::string line; std::getline(cf, line);) {
        std::istringstream is(line);
        std::string key;
        if (!(is >> key)) continue;
        std::vector<double> vals;
        for (double x; is >> x;) vals.push_back(x);
        if (key == "beam") beam_lines.push_back(vals);
        else kv[key] = vals;
    }
    auto u3 = [&](const char* k) { auto& v = kv.at(k); return uint3{(unsigned)v[0], (unsigned)v[1], (unsigned)v[2]}; };
    auto f3 = [&](const char* k) { auto& v = kv.at(k); return float3{(float)v[0], (float)v[1], (float)v[2]}; };
    auto f1 = [&](const char* k) { return (float)kv.at(k)[0]; };
    auto i1 = [&](const char* k) { return (int)kv.at(k)[0]; };
    try {
        // CONSTANTS as load_omni_header fills it (calc_header.cpp:81-161)
        CONSTANTS* constants = new CONSTANTS();
        constants->size = u3("size");
        constants->voxel = f3("voxel");
        constants->start = f3("start");
        constants->calc_bbox_start = u3("bbox_start");
        constants->calc_bbox_size = u3("bbox_size");
        constants->max_rev_size = u3("max_rev_size");
        constants->rev_latspacing = f1("rev_latspacing");
        constants->rev_longspacing = f1("rev_longspacing");
        constants->kernel_extent = f1("kernel_extent");
        constants->penumbra = f1("penumbra");
        constants->ss_factor = (uint)i1("ss_factor");
        constants->nphi = (unsigned)i1("nphi");
        constants->ntheta = (unsigned)i1("ntheta");
        constants->nradii = (unsigned)i1("nradii");
        constants->beamhard_correct = f1("beamhard_correct");
        constants->beam_count = (unsigned)beam_lines.size();
        std::strcpy(constants->beam_spec, "spec_6mv");
        std::strcpy(constants->target_structure, "PTV");
        const int nrays = (int)(constants->nphi * (constants->ntheta / 2));
        std::vector<float> theta = read_raw(dir + "/theta", (size_t)nrays), phi = read_raw(dir + "/phi", (size_t)nrays);
        constants->conv_theta_deg = theta.data();
        constants->conv_phi_deg = phi.data();
        const size_t nv = (size_t)constants->size.x * constants->size.y * constants->size.z;
        // SHM_DATA (dosecalc-beam/main.cu:200-240)
        std::vector<float> density = read_raw(dir + "/density", nv);
        std::vector<float> kernel = read_raw(dir + "/kernel", (size_t)constants->ntheta * constants->nradii);
        std::vector<float> radii = read_raw(dir + "/radii", constants->nradii);
        SHM_DATA* datavols = new SHM_DATA();
        datavols->size_data = (unsigned)nv;
        datavols->density = density.data();
        datavols->kernel_array = kernel.data();
        datavols->radial_boundary = radii.data();
        // MONO_KERNELS (read_spectrum_file, spectra.cpp:7-62)
        MONO_KERNELS* mono = new MONO_KERNELS();
        mono->nkernels = i1("nspec");
        mono->spectrum_file = "spec_6mv";
        std::vector<float> spec = read_raw(dir + "/spectrum", 4 * (size_t)mono->nkernels);
        for (int e = 0; e < mono->nkernels; e++) {
            mono->fluence[e] = spec[(size_t)e];
            mono->energy[e] = spec[(size_t)(mono->nkernels + e)];
            mono->mu_en[e] = spec[(size_t)(2 * mono->nkernels + e)];
            mono->mu[e] = spec[(size_t)(3 * mono->nkernels + e)];
        }
        // BEAMs with their fluence maps (load_omni_beam_list + fluence files, beam.cpp:516-576)
        std::vector<BEAM> beams(beam_lines.size());
        for (size_t i = 0; i < beams.size(); i++) {
            const std::vector<double>& b = beam_lines[i];
            BEAM& B = beams[i];
            B.uid = (uint)b.at(0);
            B.azimuth = (float)b[1]; B.zenith = (float)b[2]; B.coll = (float)b[3];
            B.source = float3{(float)b[4], (float)b[5], (float)b[6]};
            B.direction = float3{(float)b[7], (float)b[8], (float)b[9]};
            B.isocenter = float3{(float)b[10], (float)b[11], (float)b[12]};
            B.sad = (float)b[13];
            B.fmap_size = uint2{(unsigned)b[14], (unsigned)b[15]};
            B.beamlet_size = float2{(float)b[16], (float)b[17]};
            B.orientation_type = BEAM::ORIENT_T::ANGLE;
            B.isocenter_type = BEAM::ISO_T::PTV_CENTROID;
            B.isocenter_location = BEAM::ISO_LOC_T::IN_PTV;
            B.fluence_map = read_raw(dir + "/fmap_" + std::to_string(i), (size_t)B.fmap_size.x * B.fmap_size.y);
        }
        int gpuid_arr[1] = {0};
        initCudaConstandTex(datavols, mono, constants, nrays, 0, gpuid_arr[0]);
#ifndef CCCS_SHIM_BEAMLET
        radconvolveTexture(mono, constants, beams.data(), (int)beams.size(), nrays, 0, gpuid_arr[0]);
        std::vector<float> dose(nv);
        cuda_sum_device_float_dose(dose.data(), 1, constants);
        freeCudaTexture(1, gpuid_arr);
        std::ofstream of(out, std::ios::binary);
        of.write(reinterpret_cast<const char*>(dose.data()), (std::streamsize)(nv * sizeof(float)));
        if (!of) throw std::runtime_error("cannot write " + out);
#else
        HEADER_PATIENT header;
        header.N_beams = (ushort)beams.size();
        header.dicom_start_cm = constants->start;
        header.full_dicom_size = constants->size;
        header.voxel_size_cm = constants->voxel;
        header.rev_latspacing_cm = constants->rev_latspacing;
        header.rev_longspacing_cm = constants->rev_longspacing;
        header.bbox_start = constants->calc_bbox_start;
        header.bbox_size = constants->calc_bbox_size;
        header.penumbra_cm = constants->penumbra;
        header.sparsity_thresh = f1("threshold");
        header.nphi = (ushort)constants->nphi;
        header.ntheta = (ushort)constants->ntheta;
        header.nradii = (ushort)constants->nradii;
        header.kernel_extent_cm = constants->kernel_extent;
        header.beam_spectrum = mono->spectrum_file;
        header.target_structure = constants->target_structure;
        sparsifymanager.init(out, header, 4, 1, header.sparsity_thresh,
                             per_beamlet ? WRITE_STRATEGY::PER_BEAMLET : WRITE_STRATEGY::CENTRAL);
        sparsifymanager.activate();
        radconvolveTexture(mono, constants, beams.data(), (int)beams.size(), nrays, 0, gpuid_arr[0], 4);
        sparsifymanager.deactivate();
        freeCudaTexture(1, gpuid_arr, false);
#endif
        delete mono;
        delete datavols;
        delete constants;
    } catch (const ArraySizeError& e) {
        std::fprintf(stderr, "ArraySizeError: %s\n", e.what());
        return 3;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
