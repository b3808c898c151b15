// bh_simulate.cpp -- command-line driver with the reference's `simulate`
// interface (src/main.cpp:51-83, 185-335) on top of the C ABI of libbh.so.
//
// Same options and console lines as the reference for the single-GPU tree
// path: -f/--file, --dt, --tend, --vs, -o/--outdir, -b/--bodies, --fc,
// --theta, -r/--reference, -h/--help; --max-depth, --bucket-bits and
// --rebalance-interval are accepted and ignored (they tune the LET exchange
// and the histogram balancing, which one GPU does not need).  Differences:
//   * --fc defaults to `tree` (the reference defaults to `let`); `let` and
//     `let_direct` are refused with a message, since one GPU holds every body;
//   * --form plain|mixed|learned selects the engine's walk (default plain, the
//     reference's arithmetic bit for bit);
//   * snapshots (VTP, one part `0/`) and reference files are written with 17
//     significant digits, so a reference run and a later run compare exactly.
// Output: "Loaded N bodies", "Starting simulation (dt=.., tend=..)", one
// "Saved frame k at t=.." line per snapshot, then TOTAL_TIME (wall time of
// the step loop) and SUMMED_DIST_ERROR (sum over bodies of |x - x_ref| when
// reference/final_ref.csv exists, src/io.cpp:716-780).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/bh.h"

namespace fs = std::filesystem;

namespace {

struct Options {
    std::string file, dt = "1d", tend = "1y", vs = "10d", outdir = "sim_out", fc = "tree", form = "plain";
    int bodies = -1;
    double theta = 1.05;
    bool reference = false, help = false;
};

const char* kUsage =
    "bh_simulate -- Barnes-Hut on one MI355X (the reference's simulate interface)\n"
    "  -f, --file FILE        state-vector CSV (id,name,class,mass,x,y,z,vx,vy,vz)\n"
    "      --dt TIME          time step, e.g. 1h, 1d, 0.5y (default 1d)\n"
    "      --tend TIME        end time (default 1y)\n"
    "      --vs TIME          snapshot interval (default 10d)\n"
    "  -o, --outdir DIR       snapshot directory (default sim_out)\n"
    "  -b, --bodies N         simulate the first N bodies (default all)\n"
    "      --fc tree          force calculation (tree; let/let_direct need several ranks)\n"
    "      --theta X          opening angle (default 1.05)\n"
    "      --form F           plain | mixed | learned (default plain)\n"
    "  -r, --reference        save final positions to reference/final_ref.csv\n"
    "      --max-depth N, --bucket-bits N, --rebalance-interval N   accepted, unused\n"
    "  -h, --help             this text\n";

// Parses argv; returns an empty string or an error message.
std::string parse(int argc, char** argv, Options& o)
{
    std::map<std::string, std::string> alias = {{"-f", "--file"}, {"-o", "--outdir"}, {"-b", "--bodies"},
                                                {"-r", "--reference"}, {"-h", "--help"}};
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i], v;
        if (alias.count(a)) a = alias[a];
        const size_t eq = a.find('=');
        bool has_v = false;
        if (a.rfind("--", 0) == 0 && eq != std::string::npos) {
            v = a.substr(eq + 1);
            a = a.substr(0, eq);
            has_v = true;
        }
        if (a == "--help") { o.help = true; continue; }
        if (a == "--reference") {
            o.reference = !has_v || (v != "false" && v != "0");
            continue;
        }
        static const char* with_value[] = {"--file", "--dt", "--tend", "--vs", "--outdir", "--bodies", "--fc",
                                           "--theta", "--form", "--max-depth", "--bucket-bits",
                                           "--rebalance-interval"};
        bool known = false;
        for (const char* k : with_value) known = known || a == k;
        if (!known) return "unknown option " + a;
        if (!has_v) {
            if (i + 1 >= argc) return "option " + a + " needs a value";
            v = argv[++i];
        }
        if (a == "--file") o.file = v;
        else if (a == "--dt") o.dt = v;
        else if (a == "--tend") o.tend = v;
        else if (a == "--vs") o.vs = v;
        else if (a == "--outdir") o.outdir = v;
        else if (a == "--bodies") o.bodies = std::atoi(v.c_str());
        else if (a == "--fc") o.fc = v;
        else if (a == "--theta") o.theta = std::atof(v.c_str());
        else if (a == "--form") o.form = v;
    }
    return "";
}

void put(std::ostream& os, double v) { os << std::setprecision(17) << v; }

// One snapshot as VTK XML PolyData (src/io.cpp:84-230): points, then point
// data rank, body_id, velocity, mass, |acceleration|, then vertex cells.
bool write_vtp(const std::string& path, int64_t n, const std::vector<bh_vec3>& pos, const std::vector<bh_vec3>& vel,
               const std::vector<double>& mass, const std::vector<uint64_t>& ids, const std::vector<bh_vec3>& acc)
{
    std::ofstream os(path);
    if (!os) return false;
    os << "<?xml version=\"1.0\"?>\n<VTKFile type=\"PolyData\" version=\"0.1\">\n<PolyData>\n";
    os << "<Piece NumberOfPoints=\"" << n << "\" NumberOfVerts=\"" << n << "\">\n";
    os << "<Points>\n<DataArray type=\"Float64\" Name=\"position\" NumberOfComponents=\"3\" NumberOfTuples=\"" << n
       << "\" format=\"ascii\">\n";
    for (int64_t i = 0; i < n; ++i) {
        put(os, pos[i].x); os << ' '; put(os, pos[i].y); os << ' '; put(os, pos[i].z); os << '\n';
    }
    os << "</DataArray>\n</Points>\n<PointData>\n";
    os << "<DataArray type=\"Int32\" Name=\"rank\" NumberOfComponents=\"1\" NumberOfTuples=\"" << n
       << "\" format=\"ascii\">\n";
    for (int64_t i = 0; i < n; ++i) os << "0\n";
    os << "</DataArray>\n<DataArray type=\"UInt64\" Name=\"body_id\" NumberOfComponents=\"1\" NumberOfTuples=\"" << n
       << "\" format=\"ascii\">\n";
    for (int64_t i = 0; i < n; ++i) os << ids[i] << '\n';
    os << "</DataArray>\n<DataArray type=\"Float64\" Name=\"velocity\" NumberOfComponents=\"3\" NumberOfTuples=\"" << n
       << "\" format=\"ascii\">\n";
    for (int64_t i = 0; i < n; ++i) {
        put(os, vel[i].x); os << ' '; put(os, vel[i].y); os << ' '; put(os, vel[i].z); os << '\n';
    }
    os << "</DataArray>\n<DataArray type=\"Float64\" Name=\"mass\" NumberOfComponents=\"1\" NumberOfTuples=\"" << n
       << "\" format=\"ascii\">\n";
    for (int64_t i = 0; i < n; ++i) { put(os, mass[i]); os << '\n'; }
    os << "</DataArray>\n<DataArray type=\"Float64\" Name=\"acceleration\" NumberOfComponents=\"1\" NumberOfTuples=\""
       << n << "\" format=\"ascii\">\n";
    for (int64_t i = 0; i < n; ++i) {
        put(os, std::sqrt(acc[i].x * acc[i].x + acc[i].y * acc[i].y + acc[i].z * acc[i].z));
        os << '\n';
    }
    os << "</DataArray>\n</PointData>\n<Verts>\n<DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n";
    for (int64_t i = 1; i <= n; ++i) os << i << ' ';
    os << "\n</DataArray>\n<DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n";
    for (int64_t i = 0; i < n; ++i) os << i << ' ';
    os << "\n</DataArray>\n</Verts>\n</Piece>\n</PolyData>\n</VTKFile>\n";
    return static_cast<bool>(os);
}

// The collection file (src/io.cpp:232-298): named after the input and the run
// parameters, one DataSet per snapshot.
bool write_pvd(const std::string& path, const std::vector<std::pair<double, int>>& frames)
{
    std::ofstream os(path);
    if (!os) return false;
    os << "<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\">\n<Collection>\n";
    for (const auto& f : frames) {
        os << "<DataSet timestep=\"";
        put(os, f.first);
        os << "\" part=\"0\" file=\"0/sim." << f.second << ".vtp\"/>\n";
    }
    os << "</Collection>\n</VTKFile>\n";
    return static_cast<bool>(os);
}

int die(const std::string& msg, int code = 1)
{
    std::cerr << "Error: " << msg << "\n";
    return code;
}

}  // namespace

int main(int argc, char** argv)
{
    Options o;
    const std::string perr = parse(argc, argv, o);
    if (!perr.empty()) return die(perr + "\n" + kUsage, 2);
    if (o.help) {
        std::cout << kUsage;
        return 0;
    }
    if (o.reference) {
        o.fc = "tree";
        std::cout << "Running in reference mode. Force calculation method will be 'tree'." << std::endl;
    }
    if (o.fc != "tree")
        return die("--fc " + o.fc + " is not provided: one GPU holds every body, so only the full tree (--fc tree) "
                   "exists", 2);
    int form = BH_FORM_PLAIN;
    if (o.form == "mixed") form = BH_FORM_MIXED;
    else if (o.form == "learned") form = BH_FORM_LEARNED;
    else if (o.form != "plain") return die("unknown --form " + o.form, 2);
    double dt = 0, t_end = 0, vs = 0;
    if (bh_parse_time(o.dt.c_str(), &dt) || bh_parse_time(o.tend.c_str(), &t_end) || bh_parse_time(o.vs.c_str(), &vs))
        return die("times are a number and a unit h, d or y");
    if (!(dt > 0)) return die("--dt must be positive");
    if (o.file.empty()) return die("input file required (-f)");

    int64_t n = 0;
    if (bh_read_csv(o.file.c_str(), o.bodies, 0, &n, nullptr, nullptr, nullptr, nullptr) != BH_OK)
        return die("cannot read " + o.file);
    std::vector<bh_vec3> pos(n), vel(n), acc(n);
    std::vector<double> mass(n);
    std::vector<uint64_t> ids(n);
    int64_t got = 0;
    if (bh_read_csv(o.file.c_str(), o.bodies, n, &got, pos.data(), vel.data(), mass.data(), ids.data()) != BH_OK)
        return die("cannot read " + o.file);
    std::cout << "Loaded " << n << " bodies\n";
    const std::string ref_dir = "reference";
    std::error_code ec;
    fs::create_directories(fs::path(o.outdir) / "0", ec);
    if (o.reference) fs::create_directories(ref_dir, ec);

    bh_config cfg;
    bh_config_default(&cfg);
    cfg.form = form;
    cfg.theta = o.theta;
    bh_ctx* ctx = nullptr;
    if (bh_create(&ctx, &cfg) != BH_OK) return die(std::string("bh_create: ") + bh_last_error(nullptr));
    auto ck = [&](int rc, const char* what) {
        if (rc == BH_OK) return true;
        std::cerr << "Error: " << what << ": " << bh_last_error(ctx) << "\n";
        return false;
    };
    if (!ck(bh_set_bodies(ctx, n, pos.data(), vel.data(), mass.data(), ids.data()), "bh_set_bodies")) return 1;

    std::cout << "Starting simulation (dt=" << dt << ", tend=" << t_end << ")\n";
    std::ostringstream stem;
    stem << fs::path(o.file).stem().string() << "_dt" << o.dt << "_tend" << o.tend << "_theta" << o.theta
         << "_bodies" << o.bodies << "_fc" << o.fc << ".pvd";
    const std::string pvd = (fs::path(o.outdir) / stem.str()).string();
    std::vector<std::pair<double, int>> frames;

    // The loop of src/main.cpp:191-281: step k starts at t_k (t accumulated by
    // repeated addition); a snapshot follows step k when t_k >= next or k == 0.
    const auto t0 = std::chrono::steady_clock::now();
    auto last = t0;
    double t = 0.0, next_vis = 0.0;
    int vis = 0;
    int64_t pending = 0;   // steps queued on the device since the last snapshot
    for (int64_t step = 0; t < t_end; ++step, t += dt) {
        ++pending;
        if (!(t >= next_vis || step == 0)) continue;
        if (!ck(bh_step(ctx, pending, dt), "bh_step")) return 1;
        pending = 0;
        if (!ck(bh_get_bodies(ctx, pos.data(), vel.data(), mass.data(), ids.data(), acc.data()), "bh_get_bodies"))
            return 1;
        const std::string vtp = (fs::path(o.outdir) / "0" / ("sim." + std::to_string(vis) + ".vtp")).string();
        if (!write_vtp(vtp, n, pos, vel, mass, ids, acc)) return die("cannot write " + vtp);
        frames.emplace_back(t, vis);
        if (!write_pvd(pvd, frames)) return die("cannot write " + pvd);
        const auto now = std::chrono::steady_clock::now();
        std::cout << "Saved frame " << vis << " at t=" << t << " (Time since last frame: " << std::fixed
                  << std::setprecision(2) << std::chrono::duration<double>(now - last).count() << " s)"
                  << std::defaultfloat << std::setprecision(6) << std::endl;
        last = now;
        next_vis += vs;
        ++vis;
    }
    if (pending > 0 && !ck(bh_step(ctx, pending, dt), "bh_step")) return 1;
    if (!ck(bh_get_bodies(ctx, pos.data(), vel.data(), mass.data(), ids.data(), acc.data()), "bh_get_bodies"))
        return 1;
    const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    double summed = 0.0;
    const fs::path ref_file = fs::path(ref_dir) / "final_ref.csv";
    if (o.reference) {
        std::ofstream os(ref_file);
        os << "id,x,y,z\n";
        for (int64_t i = 0; i < n; ++i) {
            os << ids[i] << ',';
            put(os, pos[i].x); os << ','; put(os, pos[i].y); os << ','; put(os, pos[i].z); os << '\n';
        }
        if (!os) return die("cannot write " + ref_file.string());
        std::cout << "Reference positions saved to " << ref_file.string() << std::endl;
    } else if (fs::exists(ref_file)) {
        std::ifstream is(ref_file);
        std::string line;
        std::getline(is, line);
        std::vector<bh_vec3> ref(n, bh_vec3{0, 0, 0});
        while (std::getline(is, line)) {
            uint64_t id = 0;
            double x = 0, y = 0, z = 0;
            if (std::sscanf(line.c_str(), "%lu,%lf,%lf,%lf", &id, &x, &y, &z) == 4 && id < static_cast<uint64_t>(n))
                ref[id] = bh_vec3{x, y, z};
        }
        for (int64_t i = 0; i < n; ++i) {
            if (ids[i] >= static_cast<uint64_t>(n)) continue;
            const bh_vec3& p = pos[i];
            const bh_vec3& q = ref[ids[i]];
            const double dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
            summed += std::sqrt(dx * dx + dy * dy + dz * dz);
        }
    } else {
        std::cout << "Reference file not found at " << ref_file.string() << ", skipping error analysis." << std::endl;
    }
    std::cout << std::fixed << std::setprecision(4);
    std::cout << "TOTAL_TIME          " << total << std::endl;
    std::cout << "SUMMED_DIST_ERROR   " << summed << std::endl;
    bh_destroy(ctx);
    return 0;
}
