// xml.cpp -- the XML scene config of the wavesolver (ImpulseResponseParser,
// src/parser/ImpulseResponseParser.cpp:492-751): <solver>,
// <solver_control_policy type="static">, <pressure_source>/<gaussian_pressure_source>,
// <listening_point_list>, plus analytic rigid objects in <scene>:
//   <analytic_sphere center_x= center_y= center_z= radius=>
//   <analytic_ellipsoid center_x=.. semi_axis_x=.. semi_axis_y=.. semi_axis_z=..
//                       [quat_w= quat_x= quat_y= quat_z=]>
// each with any number of <gaussian_impulse_source amplitude= width= onset=/>,
// the reference's has_harmonic_source= harmonic_source_f= attributes, and
// <acc_noise_source mass= [inv_inertia= center_of_mass_x..z=]> holding
// <impact timestamp= support= impulse_x..z= position_x..z=/> elements, and the
// reference's file_kinematics= kinematics_frame_rate= keyframe animation.
// A small attribute-only XML reader replaces the vendored tinyxml.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "engine.h"

namespace fdtd {
namespace {

struct Node {
    std::string name;
    std::map<std::string, std::string> attr;
    std::vector<std::unique_ptr<Node>> kids;
    std::vector<const Node *> children(const char *n) const
    {
        std::vector<const Node *> out;
        for (auto &k : kids)
            if (k->name == n) out.push_back(k.get());
        return out;
    }
    const Node *child(const char *n) const
    {
        for (auto &k : kids)
            if (k->name == n) return k.get();
        return nullptr;
    }
};

struct Reader {
    const std::string &s;
    size_t i = 0;
    std::string err;
    explicit Reader(const std::string &src) : s(src) {}
    void ws() { while (i < s.size() && std::isspace((unsigned char)s[i])) ++i; }
    bool starts(const char *p) const { return s.compare(i, std::strlen(p), p) == 0; }
    void skip_misc()
    {
        for (;;) {
            ws();
            if (starts("<?")) { size_t e = s.find("?>", i); i = e == std::string::npos ? s.size() : e + 2; }
            else if (starts("<!--")) { size_t e = s.find("-->", i); i = e == std::string::npos ? s.size() : e + 3; }
            else if (starts("<!")) { size_t e = s.find('>', i); i = e == std::string::npos ? s.size() : e + 1; }
            else break;
        }
    }
    std::string name()
    {
        size_t b = i;
        while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '_' || s[i] == '-' || s[i] == ':' || s[i] == '.')) ++i;
        return s.substr(b, i - b);
    }
    std::unique_ptr<Node> element()
    {
        skip_misc();
        if (i >= s.size() || s[i] != '<') { err = "expected '<'"; return nullptr; }
        ++i;
        auto n = std::make_unique<Node>();
        n->name = name();
        if (n->name.empty()) { err = "empty element name"; return nullptr; }
        for (;;) {
            ws();
            if (starts("/>")) { i += 2; return n; }
            if (starts(">")) { ++i; break; }
            std::string a = name();
            if (a.empty()) { err = "bad attribute in <" + n->name + ">"; return nullptr; }
            ws();
            if (i >= s.size() || s[i] != '=') { err = "missing '=' after " + a; return nullptr; }
            ++i;
            ws();
            if (i >= s.size() || (s[i] != '"' && s[i] != '\'')) { err = "unquoted value of " + a; return nullptr; }
            const char q = s[i++];
            size_t e = s.find(q, i);
            if (e == std::string::npos) { err = "unterminated value of " + a; return nullptr; }
            n->attr[a] = s.substr(i, e - i);
            i = e + 1;
        }
        for (;;) {
            // text content is ignored
            while (i < s.size() && s[i] != '<') ++i;
            if (i >= s.size()) { err = "unterminated <" + n->name + ">"; return nullptr; }
            if (starts("<!--")) { size_t e = s.find("-->", i); i = e == std::string::npos ? s.size() : e + 3; continue; }
            if (starts("</")) {
                i += 2;
                std::string cn = name();
                ws();
                if (cn != n->name || i >= s.size() || s[i] != '>') { err = "mismatched </" + cn + "> for <" + n->name + ">"; return nullptr; }
                ++i;
                return n;
            }
            auto k = element();
            if (!k) return nullptr;
            n->kids.push_back(std::move(k));
        }
    }
};

struct ParseError {
    std::string msg;
};

const std::string *find_attr(const Node *n, const char *a)
{
    auto it = n->attr.find(a);
    return it == n->attr.end() ? nullptr : &it->second;
}
double to_real(const Node *n, const char *a, const std::string &v)
{
    char *end = nullptr;
    const double x = std::strtod(v.c_str(), &end);
    if (end == v.c_str()) throw ParseError{"<" + n->name + "> attribute " + a + " is not a number: " + v};
    return x;
}
double req_real(const Node *n, const char *a)
{
    const std::string *v = find_attr(n, a);
    if (!v) throw ParseError{"<" + n->name + "> is missing required attribute " + a};
    return to_real(n, a, *v);
}
double opt_real(const Node *n, const char *a, double def)
{
    const std::string *v = find_attr(n, a);
    return v ? to_real(n, a, *v) : def;
}
int req_int(const Node *n, const char *a) { return (int)std::lround(req_real(n, a)); }
int opt_int(const Node *n, const char *a, int def) { return (int)std::lround(opt_real(n, a, def)); }
std::string opt_str(const Node *n, const char *a, const char *def)
{
    const std::string *v = find_attr(n, a);
    return v ? *v : std::string(def);
}

}  // namespace
}  // namespace fdtd

namespace fdtd {
struct AccNoise {   // one acc_noise_source element and its impacts
    fdtd_acc_noise_source src{};
    std::vector<fdtd_impact> impacts;
};
}  // namespace fdtd

struct fdtd_parsed {
    fdtd_scene scene;
    std::vector<std::unique_ptr<fdtd::AccNoise>> acc_noise;
    std::vector<fdtd_acc_noise_source> acc_noise_src;
    std::vector<std::vector<double>> kin_frames;
    std::vector<fdtd_kinematics> kinematics;
    std::vector<fdtd_object> objects;
    std::vector<fdtd_impulse_source> impulses;
    std::vector<fdtd_pressure_source> psrc;
    std::vector<fdtd_speaker_source> spk;
    std::vector<std::vector<double>> spk_samples;
    std::vector<fdtd_plane_constraint> constraints;
    std::vector<double> listeners;
    int num_steps = 0;
    std::string output_pattern = "%s";
};

namespace {

// Channel 0 of a RIFF/WAVE file as the reference's WavReader gets it from
// libsndfile (sndgen/WavReader.hpp:96-120): integer PCM scaled to [-1, 1)
// (8-bit unsigned about 128, 16/24/32-bit signed over 2^(bits-1)), IEEE
// floats as stored.  ReadChannel reads `frames` values of the interleaved
// data, so with more than one channel only the first frames / channels
// frames carry samples and the rest are zero, as there.
bool read_wav_channel0(const std::string &path, std::vector<double> &out, double &rate, std::string &err)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) { err = "cannot open speaker file " + path; return false; }
    std::vector<unsigned char> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    auto u16 = [&](size_t o) { return (unsigned)b[o] | ((unsigned)b[o + 1] << 8); };
    auto u32 = [&](size_t o) { return (uint32_t)b[o] | ((uint32_t)b[o + 1] << 8) | ((uint32_t)b[o + 2] << 16) | ((uint32_t)b[o + 3] << 24); };
    if (b.size() < 12 || std::memcmp(&b[0], "RIFF", 4) || std::memcmp(&b[8], "WAVE", 4)) { err = path + " is not a RIFF/WAVE file"; return false; }
    unsigned fmt = 0, ch = 0, bits = 0;
    size_t data = 0, dlen = 0;
    for (size_t o = 12; o + 8 <= b.size();) {
        const uint32_t len = u32(o + 4);
        if (!std::memcmp(&b[o], "fmt ", 4) && o + 24 <= b.size()) {
            fmt = u16(o + 8); ch = u16(o + 10); rate = (double)u32(o + 12); bits = u16(o + 22);
            if (fmt == 0xFFFE && len >= 26 && o + 34 <= b.size()) fmt = u16(o + 32);   // WAVE_FORMAT_EXTENSIBLE
        } else if (!std::memcmp(&b[o], "data", 4)) {
            data = o + 8;
            dlen = std::min<size_t>(len, b.size() - data);
        }
        o += 8 + (size_t)len + (len & 1);
    }
    const unsigned bps = bits / 8;
    if (!data || !ch || !(rate > 0) || !bps || !((fmt == 1 && bps >= 1 && bps <= 4) || (fmt == 3 && (bps == 4 || bps == 8)))) {
        err = path + ": unsupported WAV layout (PCM 8/16/24/32-bit or IEEE float 32/64 expected)";
        return false;
    }
    const size_t frames = dlen / (bps * ch);
    auto value = [&](size_t item) -> double {
        const size_t o = data + item * bps;
        if (fmt == 3) {
            if (bps == 4) { float v; std::memcpy(&v, &b[o], 4); return v; }
            double v; std::memcpy(&v, &b[o], 8); return v;
        }
        if (bps == 1) return ((double)b[o] - 128.0) / 128.0;
        uint32_t raw = 0;
        for (unsigned q = 0; q < bps; ++q) raw |= (uint32_t)b[o + q] << (8 * q);
        int64_t v = raw;
        if (raw & (1u << (8 * bps - 1))) v -= (int64_t)1 << (8 * bps);
        return (double)v / (double)((int64_t)1 << (8 * bps - 1));
    };
    out.assign(frames, 0.0);
    for (size_t fr = 0; fr < frames; ++fr)
        if (fr * ch < frames) out[fr] = value(fr * ch);
    return true;
}

}  // namespace

extern "C" {

fdtd_status fdtd_parse_xml(const char *path, fdtd_parsed **out)
{
    using namespace fdtd;
    if (!path || !out) return FDTD_ERR_INVALID;
    *out = nullptr;
    std::ifstream f(path, std::ios::binary);
    if (!f) { set_last_error(std::string("cannot open ") + path); return FDTD_ERR_PARSE; }
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    Reader rd(text);
    std::unique_ptr<Node> root = rd.element();
    if (!root) { set_last_error("XML: " + rd.err); return FDTD_ERR_PARSE; }
    auto P = std::make_unique<fdtd_parsed>();
    try {
        if (root->name != "impulse_response") throw ParseError{"root element must be <impulse_response>"};
        const Node *solver = root->child("solver");
        if (!solver) throw ParseError{"missing <solver>"};
        fdtd_settings &s = P->scene.settings;
        fdtd_settings_default(&s);
        s.sound_speed = opt_real(solver, "sound_speed", 343.0);
        s.density = opt_real(solver, "density", 1.0);
        s.cell_size = req_real(solver, "cellsize");
        s.stop_time = req_real(solver, "stop_time");
        P->num_steps = opt_int(solver, "number_of_timesteps", 0);
        // output_pattern is required by the reference (ImpulseResponseParser.cpp:521);
        // here a missing one means "files in the working directory"
        P->output_pattern = opt_str(solver, "output_pattern", "%s");
        // not a reference key: the reference picks its scheme at compile time
        const std::string scheme = opt_str(solver, "scheme", "collocated");
        if (scheme == "collocated") s.scheme = FDTD_SCHEME_COLLOCATED;
        else if (scheme == "staggered") s.scheme = FDTD_SCHEME_STAGGERED;
        else throw ParseError{"scheme must be collocated or staggered: " + scheme};
        s.time_step = 1.0 / req_real(solver, "timestepfrequency");
        s.pml_width = req_real(solver, "PML_width");
        s.pml_strength = req_real(solver, "PML_strength");
        s.air_viscosity_alpha = opt_real(solver, "air_viscosity_alpha", 0.0);
        s.use_mesh = opt_int(solver, "use_mesh", 1) != 0;
        s.boundary_condition_preset = opt_int(solver, "boundary_condition_preset", 0);
        s.start_time = opt_real(solver, "fast_forward_to_event_time", 0.0);
        // time parallelization (ImpulseResponseParser.cpp:525-533)
        s.time_parallel = opt_int(solver, "time_parallel", 0) != 0;
        if (s.time_parallel) {
            s.num_time_chunks = req_int(solver, "num_time_chunks");
            s.overlap_time = opt_real(solver, "overlap_time", 6e-3);
            s.early_termination_ratio = opt_real(solver, "early_termination_ratio", 1e-3);
        }
        const std::string bh = opt_str(solver, "boundary_handling", "rasterize");
        if (bh == "rasterize") s.boundary_handling = FDTD_BOUNDARY_RASTERIZE;
        else if (bh == "fully_coupled") s.boundary_handling = FDTD_BOUNDARY_FULLY_COUPLED;
        else throw ParseError{"boundary handling type not understood: " + bh};
        const Node *pol = root->child("solver_control_policy");
        if (!pol) throw ParseError{"missing <solver_control_policy>"};
        if (opt_str(pol, "type", "") != "static") throw ParseError{"only the static solver_control_policy is supported"};
        const int res = req_int(pol, "gridresolution");
        const int div[3] = {opt_int(pol, "gridresolution_x", res), opt_int(pol, "gridresolution_y", res),
                            opt_int(pol, "gridresolution_z", res)};
        const double ctr[3] = {opt_real(pol, "domain_center_x", 0.0), opt_real(pol, "domain_center_y", 0.0),
                               opt_real(pol, "domain_center_z", 0.0)};
        // BoundingBox(cellSize, cellDivisions, center) and the ScalarField
        // division count ceil((length - EPS) / h) (ScalarField.cpp:37-46)
        for (int d = 0; d < 3; ++d) {
            const double mn = ctr[d] - s.cell_size * (double)div[d] / 2.0;
            const double mx = ctr[d] + s.cell_size * (double)div[d] / 2.0;
            s.min_bound[d] = mn;
            s.divisions[d] = (int)std::ceil(((mx - mn) - 1e-10) / s.cell_size);
        }
        if (const Node *ps = root->child("pressure_source")) {
            for (const Node *g : ps->children("gaussian_pressure_source")) {
                fdtd_pressure_source q{};
                q.width_time = req_real(g, "source_width_time");
                q.width_space = opt_real(g, "source_width_space", s.sound_speed * q.width_time);
                q.offset_time = opt_real(g, "source_offset_time", 0.0);
                const bool flip = opt_real(g, "source_sign_flip", 0.0) > 1e-10;
                q.normalize_constant = opt_real(g, "source_normalize_constant", 1.0 / std::pow(std::sqrt(2.0 * M_PI) * q.width_space, 3)) *
                                       (flip ? -1.0 : 1.0);
                q.position[0] = req_real(g, "source_position_x");
                q.position[1] = req_real(g, "source_position_y");
                q.position[2] = req_real(g, "source_position_z");
                P->psrc.push_back(q);
            }
            // ImpulseResponseParser.cpp:676-717 (static positioning)
            for (const Node *g : ps->children("speaker_pressure_source")) {
                fdtd_speaker_source q{};
                q.start_time = req_real(g, "start_time");
                q.width_space = req_real(g, "width_space");
                const std::string file = opt_str(g, "speaker_file", "");
                if (file.empty()) throw ParseError{"speaker_pressure_source needs speaker_file"};
                const std::string type = opt_str(g, "positioning_type", "static");
                if (type == "markers") throw ParseError{"speaker positioning by marker meshes is not supported"};
                if (type != "static") throw ParseError{"speaker position type not understood"};
                q.position[0] = req_real(g, "position_x");
                q.position[1] = req_real(g, "position_y");
                q.position[2] = req_real(g, "position_z");
                // the file as given, else next to the XML file
                std::vector<double> samples;
                double rate = 0;
                std::string err;
                std::string where = file;
                if (!std::ifstream(where).good() && file[0] != '/') {
                    const std::string xml(path);
                    const size_t slash = xml.find_last_of('/');
                    if (slash != std::string::npos) where = xml.substr(0, slash + 1) + file;
                }
                if (!read_wav_channel0(where, samples, rate, err)) throw ParseError{err};
                q.sample_rate = rate;
                q.n_samples = (int)samples.size();
                P->spk.push_back(q);
                P->spk_samples.push_back(std::move(samples));
            }
        }
        if (const Node *ll = root->child("listening_point_list"))
            for (const Node *l : ll->children("listening_point")) {
                P->listeners.push_back(req_real(l, "x"));
                P->listeners.push_back(req_real(l, "y"));
                P->listeners.push_back(req_real(l, "z"));
            }
        if (const Node *scn = root->child("scene")) {
            for (auto &k : scn->kids) {
                const Node *o = k.get();
                if (o->name == "constraint") {
                    // ImpulseResponseParser.cpp:434-449: direction 1, sign 1, height 0 by default
                    fdtd_plane_constraint c{};
                    c.direction = opt_int(o, "direction", 1);
                    c.sign = opt_int(o, "sign", 1);
                    c.height = opt_real(o, "height", 0.0);
                    P->constraints.push_back(c);
                    continue;
                }
                fdtd_object ob{};
                if (o->name == "analytic_sphere") {
                    ob.type = FDTD_OBJ_SPHERE;
                    const double r = req_real(o, "radius");
                    ob.semi_axes[0] = ob.semi_axes[1] = ob.semi_axes[2] = r;
                    for (int q = 0; q < 9; ++q) ob.rotation[q] = (q % 4 == 0);
                } else if (o->name == "analytic_ellipsoid") {
                    ob.type = FDTD_OBJ_ELLIPSOID;
                    ob.semi_axes[0] = req_real(o, "semi_axis_x");
                    ob.semi_axes[1] = req_real(o, "semi_axis_y");
                    ob.semi_axes[2] = req_real(o, "semi_axis_z");
                    double w = opt_real(o, "quat_w", 1.0), x = opt_real(o, "quat_x", 0.0), y = opt_real(o, "quat_y", 0.0),
                           z = opt_real(o, "quat_z", 0.0);
                    const double nq = std::sqrt(w * w + x * x + y * y + z * z);
                    w /= nq; x /= nq; y /= nq; z /= nq;
                    const double R[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                                         2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                                         2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)};
                    std::memcpy(ob.rotation, R, sizeof R);
                } else if (o->name == "rigid_object" || o->name == "rigid_sound_object") {
                    throw ParseError{"<" + o->name + "> needs mesh and distance-field files; use analytic_sphere/analytic_ellipsoid"};
                } else {
                    continue;
                }
                ob.center[0] = req_real(o, "center_x");
                ob.center[1] = req_real(o, "center_y");
                ob.center[2] = req_real(o, "center_z");
                const int oi = (int)P->objects.size();
                P->objects.push_back(ob);
                // ImpulseResponseParser.cpp:248-263: a kinematics file of rows
                // "frame tx ty tz qw qx qy qz" at kinematics_frame_rate frames
                // per second (blank lines are skipped; the reference's reader
                // would add a zero frame for one)
                const std::string kfile = opt_str(o, "file_kinematics", "N/A");
                if (kfile != "N/A") {
                    const double rate = req_real(o, "kinematics_frame_rate");
                    if (!(rate > 0)) throw ParseError{"kinematics_frame_rate must be positive"};
                    std::string where = kfile;
                    if (!std::ifstream(where).good() && kfile[0] != '/') {
                        const std::string xml(path);
                        const size_t cut = xml.find_last_of('/');
                        if (cut != std::string::npos) where = xml.substr(0, cut + 1) + kfile;
                    }
                    std::ifstream kin(where);
                    if (!kin) throw ParseError{"cannot read object kinematics file " + kfile};
                    std::vector<double> fr;
                    std::string line;
                    while (std::getline(kin, line)) {
                        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
                        std::istringstream in(line);
                        double v[8];
                        for (double &x : v)
                            if (!(in >> x)) throw ParseError{"kinematics file " + kfile + ": a row needs frame tx ty tz qw qx qy qz"};
                        fr.insert(fr.end(), v + 1, v + 8);
                    }
                    if (fr.empty()) throw ParseError{"kinematics file " + kfile + " holds no frame"};
                    fdtd_kinematics k{};
                    k.object = oi;
                    k.n_frames = (int)(fr.size() / 7);
                    k.step_size = 1.0 / rate;
                    P->kin_frames.push_back(std::move(fr));
                    P->kinematics.push_back(k);
                }
                for (const Node *g : o->children("gaussian_impulse_source")) {
                    fdtd_impulse_source is{};
                    is.object = oi;
                    is.kind = FDTD_VIB_GAUSSIAN;
                    is.amplitude = req_real(g, "amplitude");
                    is.width = req_real(g, "width");
                    is.onset = opt_real(g, "onset", 4.0 * is.width);
                    P->impulses.push_back(is);
                }
                // ImpulseResponseParser.cpp:266-275: omega = 2 pi f, phase 0,
                // unit pulsation velocity, no decay, no time range
                if (opt_real(o, "has_harmonic_source", 0.0) != 0.0) {
                    fdtd_impulse_source is{};
                    is.object = oi;
                    is.kind = FDTD_VIB_HARMONIC;
                    is.amplitude = 1.0;
                    is.omega = 2.0 * M_PI * req_real(o, "harmonic_source_f");
                    is.phase = 0.0;
                    is.decay_rate = -1.0;
                    P->impulses.push_back(is);
                }
                // ImpulseResponseParser.cpp:278-295: omega_i = 2 pi f_i, the
                // sweep's times; the amplitudes are read there but unused
                if (opt_real(o, "has_chirp_source", 0.0) != 0.0) {
                    fdtd_impulse_source is{};
                    is.object = oi;
                    is.kind = FDTD_VIB_CHIRP;
                    is.omega = 2.0 * M_PI * req_real(o, "chirp_frequency_0");
                    is.omega1 = 2.0 * M_PI * req_real(o, "chirp_frequency_1");
                    is.t0 = req_real(o, "chirp_time_0");
                    is.t1 = req_real(o, "chirp_time_1");
                    (void)req_real(o, "chirp_amplitude_0");
                    (void)req_real(o, "chirp_amplitude_1");
                    P->impulses.push_back(is);
                }
                // acceleration noise of an analytic object: its impacts listed
                // inline (the reference reads them from a rigid-body simulator's
                // impulse file against a mesh); mass, centre of mass and inverse
                // inertia default to a uniform solid of the object's shape
                for (const Node *g : o->children("acc_noise_source")) {
                    auto an = std::make_unique<AccNoise>();
                    an->src.object = oi;
                    an->src.mass = req_real(g, "mass");
                    const double m = an->src.mass, a = ob.semi_axes[0], b = ob.semi_axes[1], c = ob.semi_axes[2];
                    const double I[3] = {m * (b * b + c * c) / 5.0, m * (a * a + c * c) / 5.0, m * (a * a + b * b) / 5.0};
                    for (int d = 0; d < 9; ++d) an->src.inv_inertia[d] = d % 4 == 0 ? 1.0 / I[d / 4] : 0.0;
                    if (g->attr.count("inv_inertia")) {
                        std::istringstream in(g->attr.at("inv_inertia"));
                        for (int d = 0; d < 9; ++d)
                            if (!(in >> an->src.inv_inertia[d])) throw ParseError{"acc_noise_source inv_inertia needs 9 numbers"};
                    }
                    an->src.center_of_mass[0] = opt_real(g, "center_of_mass_x", 0.0);
                    an->src.center_of_mass[1] = opt_real(g, "center_of_mass_y", 0.0);
                    an->src.center_of_mass[2] = opt_real(g, "center_of_mass_z", 0.0);
                    for (const Node *im : g->children("impact")) {
                        fdtd_impact r{};
                        r.timestamp = req_real(im, "timestamp");
                        r.support = req_real(im, "support");
                        const char *jn[3] = {"impulse_x", "impulse_y", "impulse_z"}, *pn[3] = {"position_x", "position_y", "position_z"};
                        for (int d = 0; d < 3; ++d) { r.impulse[d] = req_real(im, jn[d]); r.position[d] = opt_real(im, pn[d], 0.0); }
                        an->impacts.push_back(r);
                    }
                    P->acc_noise.push_back(std::move(an));
                }
            }
        }
    } catch (const ParseError &e) {
        set_last_error(e.msg);
        return FDTD_ERR_PARSE;
    }
    fdtd_scene &sc = P->scene;
    sc.n_objects = (int)P->objects.size();
    sc.objects = P->objects.data();
    sc.n_impulse_sources = (int)P->impulses.size();
    sc.impulse_sources = P->impulses.data();
    for (auto &an : P->acc_noise) {
        an->src.n_impacts = (int)an->impacts.size();
        an->src.impacts = an->impacts.data();
        P->acc_noise_src.push_back(an->src);
    }
    sc.n_acc_noise_sources = (int)P->acc_noise_src.size();
    sc.acc_noise_sources = P->acc_noise_src.data();
    for (size_t q = 0; q < P->kinematics.size(); ++q) P->kinematics[q].frames = P->kin_frames[q].data();
    sc.n_kinematics = (int)P->kinematics.size();
    sc.kinematics = P->kinematics.data();
    sc.n_pressure_sources = (int)P->psrc.size();
    sc.pressure_sources = P->psrc.data();
    sc.n_listeners = (int)(P->listeners.size() / 3);
    sc.listeners = P->listeners.data();
    for (size_t q = 0; q < P->spk.size(); ++q) P->spk[q].samples = P->spk_samples[q].data();
    sc.n_speaker_sources = (int)P->spk.size();
    sc.speaker_sources = P->spk.data();
    sc.n_constraints = (int)P->constraints.size();
    sc.constraints = P->constraints.data();
    *out = P.release();
    return FDTD_OK;
}

const fdtd_scene *fdtd_parsed_scene(const fdtd_parsed *p) { return p ? &p->scene : nullptr; }
int fdtd_parsed_num_steps(const fdtd_parsed *p) { return p ? p->num_steps : 0; }
const char *fdtd_parsed_output_pattern(const fdtd_parsed *p) { return p ? p->output_pattern.c_str() : "%s"; }
void fdtd_parsed_free(fdtd_parsed *p) { delete p; }

fdtd_status fdtd_create_from_xml(const char *path, const fdtd_options *opt, fdtd_handle **out)
{
    fdtd_parsed *p = nullptr;
    fdtd_status st = fdtd_parse_xml(path, &p);
    if (st != FDTD_OK) return st;
    st = fdtd_create(&p->scene, opt, out);
    fdtd_parsed_free(p);
    return st;
}

}  // extern "C"
