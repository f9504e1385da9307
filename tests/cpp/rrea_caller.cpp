// A compiled caller of include/rrea_hip.h laid out as a WarpX overlay calls it
// (the integration guide's steps 2-4): tables once from the loaded bundle's
// columns, fields and density for the step, then Apply's species passes in
// the reference's order -- electrons, photons, positrons
// (RreaParticleInteraction.cpp:57-59) -- over device SoA tiles whose pre-push
// snapshot (rrea_prev_*) and live arrays are different buffers, as
// CaptureRreaPositionsBeforePush leaves them (RreaWarpXCoupling.cpp:151-208),
// and the newborn residual advance of DrainAndCommitSecondaries over the three
// passes' ordered records (RreaSecondaryCommit.cpp:338-468).
//
//   rrea_caller <input dir> <output dir>
//
// The input directory holds one raw little-endian file per array (<name>.f8
// for doubles, <name>.i8 for int64, <name>.i4 for int32) and meta.txt with
// "key value" lines; the outputs are written the same way.  Exit status 0 on
// success; any failing call prints its name and status and exits 1.

#include "rrea_hip.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

namespace {

std::string g_in, g_out;

template <class V>
std::vector<V> load(std::string const& name, char const* suffix)
{
    std::ifstream f(g_in + "/" + name + suffix, std::ios::binary | std::ios::ate);
    if (!f) { std::fprintf(stderr, "missing input %s%s\n", name.c_str(), suffix); std::exit(1); }
    std::streamsize const bytes = f.tellg();
    f.seekg(0);
    std::vector<V> v(static_cast<size_t>(bytes) / sizeof(V));
    f.read(reinterpret_cast<char*>(v.data()), bytes);
    return v;
}

std::vector<double> f8(std::string const& n) { return load<double>(n, ".f8"); }
std::vector<int64_t> i8(std::string const& n) { return load<int64_t>(n, ".i8"); }
std::vector<int32_t> i4(std::string const& n) { return load<int32_t>(n, ".i4"); }

template <class V>
void save(std::string const& name, char const* suffix, V const* p, size_t n)
{
    std::ofstream f(g_out + "/" + name + suffix, std::ios::binary);
    f.write(reinterpret_cast<char const*>(p), static_cast<std::streamsize>(n * sizeof(V)));
}

std::map<std::string, double> meta()
{
    std::map<std::string, double> m;
    std::ifstream f(g_in + "/meta.txt");
    std::string k;
    double v;
    while (f >> k >> v) m[k] = v;
    return m;
}

void check(rrea_status s, char const* what)
{
    if (s != RREA_OK) { std::fprintf(stderr, "%s failed: %d\n", what, static_cast<int>(s)); std::exit(1); }
}

void hip(hipError_t e, char const* what)
{
    if (e != hipSuccess) { std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e)); std::exit(1); }
}

// device copy of a host array (kept for the run)
template <class V>
V* upload(std::vector<V> const& h)
{
    V* d = nullptr;
    hip(hipMalloc(&d, (h.empty() ? 1 : h.size()) * sizeof(V)), "hipMalloc");
    if (!h.empty()) hip(hipMemcpy(d, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice), "hipMemcpy");
    return d;
}

template <class V>
V* zeros(size_t n)
{
    V* d = nullptr;
    hip(hipMalloc(&d, (n ? n : 1) * sizeof(V)), "hipMalloc");
    hip(hipMemset(d, 0, (n ? n : 1) * sizeof(V)), "hipMemset");
    return d;
}

template <class V>
std::vector<V> download(V const* d, size_t n)
{
    std::vector<V> h(n);
    if (n) hip(hipMemcpy(h.data(), d, n * sizeof(V), hipMemcpyDeviceToHost), "hipMemcpy");
    return h;
}

// The columns of the loaded bundle, in the header's slot order (step 2).
struct TableColumns {
    std::vector<std::vector<double>> d;
    std::vector<std::vector<int64_t>> i;
    double const* D(std::string const& n) { d.push_back(f8(n)); return d.back().data(); }
    int64_t const* I(std::string const& n) { i.push_back(i8(n)); return i.back().data(); }
    int64_t size_of_last_d() const { return static_cast<int64_t>(d.back().size()); }
};

rrea_tables_t make_tables(TableColumns& c)
{
    static char const* const rate_keys[15] = {"e_col", "e_soft", "e_moller", "e_brems", "e_eltr",
                                              "p_col", "p_soft", "p_bhabha", "p_brems", "p_eltr", "p_annih",
                                              "g_compton", "g_pe", "g_pairn", "g_pairt"};
    rrea_table_image im{};
    for (int t = 0; t < 15; ++t) {
        std::string const k = rate_keys[t];
        im.rate[t].x = c.D(k + "_x");
        im.rate[t].n = c.size_of_last_d();
        im.rate[t].y = c.D(k + "_y");
        im.rate[t].mode = static_cast<int32_t>(i8(k + "_mode").at(0));
    }
    auto cdf = [&](rrea_cdf_groups& g, std::string const& k) {
        g.primary = c.D(k + "_primary");
        g.n_groups = c.size_of_last_d();
        g.start = c.I(k + "_start");
        g.u = c.D(k + "_u");
        g.v = c.D(k + "_v");
    };
    cdf(im.moller, "moller");
    cdf(im.bhabha, "bhabha");
    cdf(im.electron_brems, "ebrems");
    cdf(im.positron_brems, "pbrems");
    auto angle = [&](rrea_angle_table& a, std::string const& k) {
        a.e = c.D(k + "_e"); a.n_e = c.size_of_last_d();
        a.n = c.D(k + "_n"); a.n_n = c.size_of_last_d();
        a.u = c.D(k + "_u"); a.n_u = c.size_of_last_d();
        a.omc = c.D(k + "_omc");
    };
    angle(im.gs_electron, "gs_e");
    angle(im.gs_positron, "gs_p");
    angle(im.wentzel_electron, "wz_e");
    angle(im.wentzel_positron, "wz_p");
    auto elastic = [&](rrea_elastic_table& e, std::string const& k) {
        e.e = c.D(k + "_e"); e.n = c.size_of_last_d();
        e.mfp_m_stp = c.D(k + "_mfp");
        e.screening_eta = c.D(k + "_eta");
    };
    elastic(im.elastic_electron, "el_e");
    elastic(im.elastic_positron, "el_p");
    std::vector<double> const s = f8("scalars");
    im.hard_moller_secondary_threshold_eV = s.at(0);
    im.brems_absolute_cutoff_eV = s.at(1);
    im.reference_density_kg_m3 = s.at(2);
    im.charged_energy_range_eV[0] = s.at(3);
    im.charged_energy_range_eV[1] = s.at(4);
    im.photon_energy_range_eV[0] = s.at(5);
    im.photon_energy_range_eV[1] = s.at(6);
    im.gs_few_collision_threshold_n = s.at(7);
    im.gs_isotropic_above_path_n = s.at(8);
    im.gs_wentzel_transition_energy_eV = s.at(9);
    im.photoelectric.primary = c.D("pe_primary");
    im.photoelectric.n_groups = c.size_of_last_d();
    im.photoelectric.start = c.I("pe_start");
    im.photoelectric.cdf_u = c.D("pe_u");
    im.photoelectric.binding_eV = c.D("pe_binding");
    im.pair.channel = c.I("pair_block_channel");
    im.pair.energy_eV = c.D("pair_block_energy");
    im.pair.n_blocks = c.size_of_last_d();
    im.pair.start = c.I("pair_block_start");
    im.pair.fraction = c.D("pair_frac");
    im.pair.direction = c.D("pair_dir");
    rrea_tables_t T = nullptr;
    check(rrea_tables_create(&im, &T), "rrea_tables_create");
    return T;
}

// One species' device SoA: the pre-push snapshot and the live arrays apart.
struct Species {
    std::string name;
    int64_t n = 0;
    double *prev_r, *prev_z, *prev_theta, *prev_ux = nullptr, *prev_uy = nullptr, *prev_uz = nullptr;
    double *w, *r, *z, *theta, *ux, *uy, *uz;
    int64_t* id;
    int32_t* cpu;
    // the step's outputs
    uint8_t* status;
    double* ion_pairs;
    int32_t* substeps;
    rrea_secondary* sec;
    int64_t* sec_count;
    int64_t sec_cap;

    Species(std::string const& nm, bool photon, int64_t cap) : name(nm), sec_cap(cap)
    {
        std::vector<double> const r0 = f8(nm + "_prev_r");
        n = static_cast<int64_t>(r0.size());
        prev_r = upload(r0);
        prev_z = upload(f8(nm + "_prev_z"));
        prev_theta = upload(f8(nm + "_prev_theta"));
        if (!photon) {
            prev_ux = upload(f8(nm + "_prev_ux"));
            prev_uy = upload(f8(nm + "_prev_uy"));
            prev_uz = upload(f8(nm + "_prev_uz"));
        }
        w = upload(f8(nm + "_w"));
        id = upload(i8(nm + "_id"));
        cpu = upload(i4(nm + "_cpu"));
        r = upload(f8(nm + "_r"));
        z = upload(f8(nm + "_z"));
        theta = upload(f8(nm + "_theta"));
        ux = upload(f8(nm + "_ux"));
        uy = upload(f8(nm + "_uy"));
        uz = upload(f8(nm + "_uz"));
        status = zeros<uint8_t>(n);
        ion_pairs = zeros<double>(n);
        substeps = zeros<int32_t>(n);
        sec = zeros<rrea_secondary>(cap);
        sec_count = zeros<int64_t>(1);
    }

    rrea_tile tile() const
    {
        return rrea_tile{n, prev_r, prev_z, prev_theta, prev_ux, prev_uy, prev_uz, w, id, cpu, r, z, theta, ux, uy, uz};
    }

    void write_back() const
    {
        auto const N = static_cast<size_t>(n);
        for (auto const& [k, p] : {std::pair<char const*, double*>{"_r", r}, {"_z", z}, {"_theta", theta},
                                   {"_ux", ux}, {"_uy", uy}, {"_uz", uz}, {"_prev_r", prev_r}, {"_prev_z", prev_z},
                                   {"_prev_theta", prev_theta}, {"_ion_pairs", ion_pairs}}) {
            std::vector<double> const h = download(p, N);
            save(name + k, ".f8", h.data(), N);
        }
        std::vector<int64_t> const ids = download(id, N);
        save(name + "_id", ".i8", ids.data(), N);
        std::vector<uint8_t> const st = download(status, N);
        std::vector<int32_t> st4(st.begin(), st.end());
        save(name + "_status", ".i4", st4.data(), N);
        std::vector<int32_t> const sub = download(substeps, N);
        save(name + "_substeps", ".i4", sub.data(), N);
        int64_t const c = download(sec_count, 1).at(0);
        size_t const m = static_cast<size_t>(c < sec_cap ? c : sec_cap);
        std::vector<rrea_secondary> const s = download(sec, m);
        save(name + "_secondaries", ".bin", s.data(), m);
    }
};

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 3) { std::fprintf(stderr, "usage: rrea_caller <input dir> <output dir>\n"); return 2; }
    g_in = argv[1];
    g_out = argv[2];
    std::map<std::string, double> m = meta();
    hipStream_t stream = nullptr;
    hip(hipStreamCreate(&stream), "hipStreamCreate");

    // step 2: tables, once after RreaInteractionTables::Load
    TableColumns columns;
    rrea_tables_t const T = make_tables(columns);
    int32_t const nr = static_cast<int32_t>(m.at("nr")), nz = static_cast<int32_t>(m.at("nz"));
    rrea_rz_geometry const g{nr, nz, m.at("dr"), m.at("dz"), 0.0, m.at("z_lo")};
    rrea_fields_t F = nullptr;
    check(rrea_fields_create(&g, &F), "rrea_fields_create");

    // step 3: fields and density, every step before transport
    std::vector<double> const er = f8("er"), ez = f8("ez"), amb_er = f8("amb_er"), amb_ez = f8("amb_ez");
    std::vector<double> const bt = f8("btheta");
    check(rrea_fields_update(F, er.data(), ez.data(), amb_er.data(), amb_ez.data(), bt.data(), 0, stream),
          "rrea_fields_update");
    if (m.count("altitude_at_z0_m")) {
        std::vector<double> const alt = f8("profile_alt"), rho = f8("profile_rho");
        check(rrea_density_set_profile(F, alt.data(), rho.data(), static_cast<int64_t>(alt.size()),
                                       m.at("altitude_at_z0_m"), m.at("reference_density")),
              "rrea_density_set_profile");
    } else {
        check(rrea_density_set_constant(F, m.at("density_ratio")), "rrea_density_set_constant");
    }

    // step 4: Apply's species passes
    rrea_step_config cfg{};
    cfg.rng_seed = static_cast<uint64_t>(m.at("rng_seed"));
    cfg.step = static_cast<int64_t>(m.at("step"));
    cfg.dt_s = m.at("dt_s");
    cfg.low_energy_cutoff_eV = m.at("low_energy_cutoff_eV");
    cfg.photon_cutoff_eV = m.at("photon_cutoff_eV");
    cfg.max_optical_depth_per_substep = m.at("max_optical_depth_per_substep");
    cfg.max_subcycles = static_cast<uint64_t>(m.at("max_subcycles"));
    cfg.strict_transport_guards = static_cast<int32_t>(m.at("strict_transport_guards"));
    cfg.disable_hard_moller = 0;
    cfg.max_gs_collisions_per_substep = m.at("max_gs_collisions_per_substep");
    cfg.gs_total_collision_cap = m.at("gs_total_collision_cap");
    cfg.max_field_impulse_fraction = m.at("max_field_impulse_fraction");
    cfg.field_impulse_floor_eV = m.at("field_impulse_floor_eV");
    int64_t const cap = static_cast<int64_t>(m.at("secondary_capacity"));
    Species e("electron", false, cap), p("positron", false, cap), ph("photon", true, cap);
    // the coupling's net-chord grid, material sources off, the step's tallies
    size_t const nrp = static_cast<size_t>(nr) + 2, nzp = static_cast<size_t>(nz) + 4;
    double* jr = zeros<double>((nrp + 1) * nzp);
    double* jz = zeros<double>(nrp * (nzp + 1));
    double* rho = zeros<double>(nrp * nzp);
    double* tallies = zeros<double>(RREA_TALLY_COUNT);
    int32_t* error = zeros<int32_t>(2);
    auto outputs = [&](Species const& s) {
        return rrea_step_out{s.status, s.ion_pairs, s.substeps, s.sec, s.sec_cap, s.sec_count, jr, jz, rho, tallies,
                             error, 2, nullptr};
    };
    {
        rrea_tile const t = e.tile();
        rrea_step_out const o = outputs(e);
        check(rrea_transport_step_electrons(T, F, &t, &cfg, &o, stream), "rrea_transport_step_electrons");
    }
    {
        rrea_tile const t = ph.tile();
        rrea_step_out const o = outputs(ph);
        check(rrea_transport_step_photons(T, F, &t, &cfg, &o, stream), "rrea_transport_step_photons");
    }
    {
        rrea_tile const t = p.tile();
        rrea_step_out const o = outputs(p);
        check(rrea_transport_step_positrons(T, F, &t, &cfg, &o, stream), "rrea_transport_step_positrons");
    }
    // DrainAndCommitSecondaries: the passes' ordered records, electrons, photons, positrons
    int64_t const fifo = static_cast<int64_t>(m.at("fifo_capacity"));
    rrea_secondary const* staged[3] = {e.sec, ph.sec, p.sec};
    int64_t const* counts[3] = {e.sec_count, ph.sec_count, p.sec_count};
    int64_t const caps[3] = {cap, cap, cap};
    rrea_secondary* work = zeros<rrea_secondary>(fifo);
    rrea_secondary* survivors = zeros<rrea_secondary>(fifo);
    int64_t* work_count = zeros<int64_t>(1);
    int64_t* survivor_count = zeros<int64_t>(1);
    uint8_t* d_status = zeros<uint8_t>(fifo);
    double* d_ion = zeros<double>(fifo);
    int32_t* d_sub = zeros<int32_t>(fifo);
    rrea_drain_io const io{work, fifo, 0, work_count, d_status, d_ion, d_sub, survivors, fifo, survivor_count,
                           jr, jz, rho, tallies, error, nullptr};
    check(rrea_newborn_drain_staged(T, F, &cfg, staged, counts, caps, 3, &io,
                                    static_cast<int32_t>(m.at("max_generations")), stream),
          "rrea_newborn_drain_staged");
    hip(hipStreamSynchronize(stream), "hipStreamSynchronize");

    for (Species const* s : {&e, &ph, &p}) s->write_back();
    int64_t const ns = download(survivor_count, 1).at(0);
    std::vector<rrea_secondary> const sv = download(survivors, static_cast<size_t>(ns));
    save("survivors", ".bin", sv.data(), sv.size());
    int64_t const nw = download(work_count, 1).at(0);
    std::vector<rrea_secondary> const wk = download(work, static_cast<size_t>(nw));
    save("work", ".bin", wk.data(), wk.size());
    std::vector<double> const t = download(tallies, RREA_TALLY_COUNT);
    save("tallies", ".f8", t.data(), t.size());
    std::vector<int32_t> const err = download(error, 2);
    save("error", ".i4", err.data(), err.size());
    for (auto const& [k, ptr, n] : {std::tuple<char const*, double*, size_t>{"jr", jr, (nrp + 1) * nzp},
                                    {"jz", jz, nrp * (nzp + 1)}, {"rho", rho, nrp * nzp}}) {
        std::vector<double> const h = download(ptr, n);
        save(k, ".f8", h.data(), n);
    }
    check(rrea_fields_destroy(F), "rrea_fields_destroy");
    check(rrea_tables_destroy(T), "rrea_tables_destroy");
    std::printf("rrea_caller: %lld e-, %lld gamma, %lld e+, %lld FIFO entries, %lld survivors, error %d\n",
                static_cast<long long>(e.n), static_cast<long long>(ph.n), static_cast<long long>(p.n),
                static_cast<long long>(nw), static_cast<long long>(ns), err.at(0));
    return 0;
}
