// vx3mi_run -- command-line drop-in for voxcraft-sim / vx3_node_worker
// (src/Executables/voxcraft-sim.cpp:25-132, vx3_node_worker.cu:31-141):
//
//     vx3mi_run -i <input_dir | task.vxt> [-o <report.xml>] [-f] [--fp64] [--max-steps N]
//
// <input_dir> holds base.vxa and the robots as *.vxd (or a worker task file
// names the base VXA, the input directory and the VXD files); every VXD is merged into
// the base (replace= paths) and all robots run as one batch on the GPU.  The
// .history text (RecordStepSize > 0) and the status lines go to stdout robot
// by robot; the report is the reference's report XML, results sorted by
// fitness with diverged (NaN) ones last.  Only the C ABI of libvx3mi is used.
#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "vx3mi.h"

namespace {

const char *kGreen = "\033[0;32m", *kBlue = "\033[0;34m", *kRed = "\033[1;31m\n", *kReset = "\033[0m";

void usage() {
    std::printf(
        "vx3mi_run: runs a generation of VXA/VXD robots on the GPU.\n"
        "Usage: vx3mi_run -i <input_dir> [-o <report.xml>] [-f] [--fp64] [--max-steps N] [--mode auto|tiled|persistent|flat|plain]\n"
        "  -i  directory with base.vxa and *.vxd files, or a vx3_node_worker task file (.vxt)\n"
        "  -o  report file (XML)\n"
        "  -f  overwrite the report if it exists\n");
}

bool is_file(const std::string &p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}
bool is_dir(const std::string &p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

std::string xml_unescape(const std::string &s) {
    static const char *ent[5][2] = {{"&amp;", "&"}, {"&lt;", "<"}, {"&gt;", ">"}, {"&quot;", "\""}, {"&apos;", "'"}};
    std::string o;
    for (size_t i = 0; i < s.size();) {
        bool done = false;
        for (auto &e : ent)
            if (s.compare(i, std::strlen(e[0]), e[0]) == 0) { o += e[1]; i += std::strlen(e[0]); done = true; break; }
        if (!done) o += s[i++];
    }
    return o;
}

// the text of the first <tag>...</tag> in [from, to) of s; npos-safe
bool element(const std::string &s, const std::string &tag, size_t from, size_t to, std::string &out, size_t *end = nullptr) {
    const std::string open = "<" + tag + ">", close = "</" + tag + ">";
    const size_t a = s.find(open, from);
    if (a == std::string::npos || a >= to) return false;
    const size_t b = s.find(close, a + open.size());
    if (b == std::string::npos || b > to) return false;
    out = s.substr(a + open.size(), b - a - open.size());
    if (end) *end = b + close.size();
    return true;
}

// A vx3_node_worker task file (.vxt: Boost ptree XML that voxcraft-sim writes,
// src/Executables/voxcraft-sim.cpp:101-110, and the worker reads,
// vx3_node_worker.cu:63-75): the base VXA path, the input directory and the VXD
// file names under vxd/f, in order.
bool read_task(const std::string &path, std::string &base, std::string &dir, std::vector<std::string> &files) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::string s;
    char buf[4096];
    for (size_t n; (n = std::fread(buf, 1, sizeof(buf), f)) > 0;) s.append(buf, n);
    std::fclose(f);
    std::string vxd;
    if (!element(s, "vxa", 0, s.size(), base) || !element(s, "input_dir", 0, s.size(), dir) ||
        !element(s, "vxd", 0, s.size(), vxd))
        return false;
    base = xml_unescape(base);
    dir = xml_unescape(dir);
    std::string name;
    for (size_t at = 0, end = 0; element(vxd, "f", at, vxd.size(), name, &end); at = end) files.push_back(xml_unescape(name));
    return true;
}

std::string xml_escape(const std::string &s) {
    std::string o;
    for (char c : s) {
        if (c == '&') o += "&amp;";
        else if (c == '<') o += "&lt;";
        else if (c == '>') o += "&gt;";
        else o += c;
    }
    return o;
}

std::string num(double v) {
    char b[64];
    std::snprintf(b, sizeof(b), "%.17g", v);
    return b;
}

// boost::filesystem path::filename(): the last component, "." after a trailing separator
std::string path_filename(const std::string &p) {
    if (p.empty()) return "";
    if (p.back() == '/') return ".";
    const size_t k = p.find_last_of('/');
    return k == std::string::npos ? p : p.substr(k + 1);
}

struct Sim {
    std::string file, name;
    vx3mi_model *model = nullptr;
    vx3mi_result res{};
    std::string history;
    std::vector<double> finalPos, initPos;
    std::vector<int> mats;
    int nv = 0;
};

void on_history(void *user, int robot, const char *text, size_t len) {
    std::vector<Sim> &sims = *static_cast<std::vector<Sim> *>(user);
    sims[robot].history.append(text, len);
}

std::string to_string_f(double v) {  // std::to_string(double)
    char b[64];
    std::snprintf(b, sizeof(b), "%f", v);
    return b;
}

}  // namespace

int main(int argc, char **argv) {
    std::string input, output, modeName = "auto";
    bool force = false;
    int precision = VX3MI_FP32;
    long long maxSteps = 0;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if ((a == "-i" || a == "--input") && i + 1 < argc) input = argv[++i];
        else if ((a == "-o" || a == "--output") && i + 1 < argc) output = argv[++i];
        else if (a == "-f" || a == "--force") force = true;
        else if (a == "--fp64") precision = VX3MI_FP64;
        else if (a == "--max-steps" && i + 1 < argc) maxSteps = std::atoll(argv[++i]);
        else if (a == "--mode" && i + 1 < argc) modeName = argv[++i];
        else if (a == "-l" || a == "--locally") continue;  // the reference's default; accepted
        else { usage(); return 1; }
    }
    if (input.empty()) { usage(); return 1; }
    if (!output.empty() && is_file(output) && !force) {
        std::printf("Error: output file exists.\n\n");
        usage();
        return 1;
    }
    std::string dir = input, base, reportDir = input;
    std::vector<std::string> files;
    const bool task = is_file(input);
    if (task) {
        if (!read_task(input, base, dir, files)) {
            std::printf("Error: %s is not a task file (vxa, input_dir, vxd/f).\n\n", input.c_str());
            usage();
            return 1;
        }
        reportDir = dir;
        if (!dir.empty() && dir.back() != '/') dir += '/';
        if (!is_file(base)) { std::printf("Error: base VXA %s not found.\n", base.c_str()); return 1; }
    } else {
        if (!is_dir(input)) { std::printf("Error: input directory not found.\n\n"); usage(); return 1; }
        if (dir.back() != '/') dir += '/';
        base = dir + "base.vxa";
        if (!is_file(base)) { std::printf("No base.vxa found in input directory.\n\n"); usage(); return 1; }
    }
    int mode = VX3MI_MODE_AUTO;
    if (modeName == "tiled") mode = VX3MI_MODE_TILED;
    else if (modeName == "persistent") mode = VX3MI_MODE_PERSISTENT;
    else if (modeName == "flat") mode = VX3MI_MODE_FLAT;
    else if (modeName == "plain") mode = VX3MI_MODE_PLAIN;

    if (!task) {  // a task file keeps its own order
        if (DIR *d = opendir(input.c_str())) {
            while (dirent *e = readdir(d)) {
                std::string f = e->d_name;
                if (f.size() < 4) continue;
                std::string ext = f.substr(f.size() - 4);
                for (char &c : ext) c = (char)std::tolower((unsigned char)c);
                if (ext == ".vxd" && is_file(dir + f)) files.push_back(f);
            }
            closedir(d);
        }
        std::sort(files.begin(), files.end());
    }
    if (files.empty()) { std::printf("No .vxd files found in input directory.\n"); return 1; }

    std::vector<Sim> sims(files.size());
    std::printf("=== set device to 0 for %zu simulations ===\n", files.size());
    for (size_t k = 0; k < files.size(); k++) {
        Sim &s = sims[k];
        s.file = files[k];
        s.name = s.file.substr(0, s.file.find('.'));
        if (vx3mi_model_from_vxa(base.c_str(), (dir + s.file).c_str(), &s.model) != VX3MI_OK) {
            std::printf("%s%s: %s%s\n", kRed, s.file.c_str(), vx3mi_last_error(), kReset);
            return 1;
        }
        std::printf("%s\n", vx3mi_model_raw_print(s.model));
        int nl, ns;
        vx3mi_model_counts(s.model, &s.nv, &nl, &ns);
    }
    std::vector<vx3mi_model *> models;
    for (Sim &s : sims) models.push_back(s.model);
    vx3mi_batch *batch = nullptr;
    if (vx3mi_batch_create(models.data(), (int)models.size(), precision, &batch) != VX3MI_OK) {
        std::printf("%s%s%s\n", kRed, vx3mi_last_error(), kReset);
        return 1;
    }
    std::vector<vx3mi_result> results(sims.size());
    std::vector<double *> finals(sims.size());
    for (size_t k = 0; k < sims.size(); k++) {
        sims[k].finalPos.resize((size_t)3 * sims[k].nv);
        finals[k] = sims[k].finalPos.data();
    }
    const int rc = vx3mi_run(batch, models.data(), (int)models.size(), maxSteps, mode, on_history, &sims, results.data(),
                             finals.data());
    if (rc != VX3MI_OK) {
        std::printf("%s%s%s\n", kRed, vx3mi_last_error(), kReset);
        return 1;
    }
    for (size_t k = 0; k < sims.size(); k++) {
        Sim &s = sims[k];
        s.res = results[k];
        vx3mi_settings st;
        vx3mi_model_settings(s.model, &st);
        std::printf("%s0) Simulation %zu runs: %s.\n%s", kGreen, k, s.file.c_str(), kReset);
        // the history header precedes the real_stepsize line, frames follow it
        const size_t hdrEnd = st.record_step_size ? s.history.find("</voxel_size>\n") : std::string::npos;
        if (hdrEnd != std::string::npos) std::fwrite(s.history.data(), 1, hdrEnd + 14, stdout);
        std::printf("real_stepsize: %d ; recommendedTimeStep %f; d_v3->DtFrac %f . \n", s.res.real_stepsize,
                    s.res.recommended_dt, st.dt_frac);
        if (hdrEnd != std::string::npos) std::fwrite(s.history.data() + hdrEnd + 14, 1, s.history.size() - hdrEnd - 14, stdout);
        if (s.res.diverged) std::printf("%s\n0) Simulation %zu Diverged: %s.\n%s", kRed, k, s.file.c_str(), kReset);
        std::printf("%s0) Simulation %zu ends: %s Time: %f, angleSampleTimes: %d.\n%s", kBlue, k, s.file.c_str(),
                    s.res.current_time, s.res.angle_sample_times, kReset);
    }
    std::fflush(stdout);

    if (!output.empty()) {
        std::vector<size_t> order(sims.size());
        for (size_t k = 0; k < order.size(); k++) order[k] = k;
        // VX3_SimulationResult::compareFitnessScore: descending, NaN last
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
            const double fa = sims[a].res.fitness_score, fb = sims[b].res.fitness_score;
            if (std::isnan(fb)) return !std::isnan(fa);
            if (std::isnan(fa)) return false;
            return fa > fb;
        });
        std::string x = "<?xml version=\"1.0\" encoding=\"utf-8\"?>\n<report>";
        x += "<inputdir>" + xml_escape(path_filename(reportDir)) + "</inputdir>";
        const Sim &best = sims[order[0]];
        x += "<bestfit><filename>" + xml_escape(best.file) + "</filename><fitness_score>" + num(best.res.fitness_score) +
             "</fitness_score></bestfit><detail>";
        std::vector<std::string> seen;
        for (size_t k : order) {
            const Sim &s = sims[k];
            if (std::find(seen.begin(), seen.end(), s.name) != seen.end()) continue;  // ptree put overwrites: keep one
            seen.push_back(s.name);
            const vx3mi_result &r = s.res;
            std::string e = "<" + s.name + ">";
            e += "<currentTime>" + num(r.current_time) + "</currentTime>";
            e += "<fitness_score>" + num(r.fitness_score) + "</fitness_score>";
            e += "<num_voxel>" + std::to_string(r.num_voxel) + "</num_voxel>";
            e += "<num_measured_voxel>" + std::to_string(r.num_measured_voxel) + "</num_measured_voxel>";
            e += "<voxSize>" + num(r.vox_size) + "</voxSize>";
            e += "<numClosePairs>" + std::to_string(r.num_close_pairs) + "</numClosePairs>";
            e += "<initialCenterOfMass><x>" + num(r.initial_com[0]) + "</x><y>" + num(r.initial_com[1]) + "</y><z>" +
                 num(r.initial_com[2]) + "</z></initialCenterOfMass>";
            e += "<currentCenterOfMass><x>" + num(r.current_com[0]) + "</x><y>" + num(r.current_com[1]) + "</y><z>" +
                 num(r.current_com[2]) + "</z></currentCenterOfMass>";
            e += "<total_distance_of_all_voxels>" + num(r.total_distance_of_all_voxels) + "</total_distance_of_all_voxels>";
            if (vx3mi_model_save_positions(s.model)) {
                std::vector<int> lat((size_t)3 * s.nv);
                vx3mi_model_topology(s.model, lat.data(), nullptr, nullptr, nullptr);
                vx3mi_settings st;
                vx3mi_model_settings(s.model, &st);
                std::string ip, fp, mt;
                for (int v = 0; v < s.nv; v++) {
                    for (int c = 0; c < 3; c++) ip += to_string_f(lat[3 * v + c] * st.lattice_dim) + (c < 2 ? "," : ";");
                    for (int c = 0; c < 3; c++) fp += to_string_f(s.finalPos[3 * v + c]) + (c < 2 ? "," : ";");
                }
                std::vector<int> pal((size_t)s.nv);
                vx3mi_model_palette_index(s.model, pal.data());
                for (int v = 0; v < s.nv; v++) mt += std::to_string(pal[v]) + ";";
                e += "<init_pos>" + ip + "</init_pos><pos>" + fp + "</pos><mats>" + mt + "</mats>";
            }
            e += "</" + s.name + ">";
            x += e;
        }
        x += "</detail></report>\n";
        FILE *f = std::fopen(output.c_str(), "wb");
        if (!f) { std::printf("ERROR: cannot write %s\n", output.c_str()); return 1; }
        std::fwrite(x.data(), 1, x.size(), f);
        std::fclose(f);
    }
    vx3mi_batch_destroy(batch);
    for (Sim &s : sims) vx3mi_model_free(s.model);
    return 0;
}
