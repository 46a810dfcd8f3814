// Derived-type restart files of the reference: reader and writer (see fesomtra_restart.hpp).
#include "fesomtra_restart.hpp"

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <sys/stat.h>

namespace fesomtra {
namespace {

// Item sequences of the write procedures.  "i" int32, "r" real64, "l" logical, "c"
// character(20); "I<k>" / "R<k>" an allocatable int32 / real64 array of rank k; "S<n>" a fixed
// int32 array of n elements (write1d_int_static: extent, then elements).
const char *kMesh =
    "i:nod2D r:ocean_area r:ocean_areawithcav i:edge2D i:edge2D_in i:elem2D I2:elem2D_nodes I2:edges I2:edge_tri "
    "I2:elem_edges R1:elem_area R2:edge_dxdy R2:edge_cross_dxdy R1:elem_cos R1:metric_factor I2:elem_neighbors "
    "I2:nod_in_elem2D R2:x_corners R2:y_corners I1:nod_in_elem2D_num R1:depth R2:gradient_vec R2:gradient_sca "
    "I1:bc_index_nod2D i:nl R1:zbar R1:Z R1:elem_depth I1:ulevels I1:ulevels_nod2D I1:ulevels_nod2D_max I1:nlevels "
    "I1:nlevels_nod2D I1:nlevels_nod2D_min R2:area R2:area_inv R2:areasvol R2:areasvol_inv R1:mesh_resolution "
    "I1:cavity_flag_n I1:cavity_flag_e R1:cavity_depth R2:cavity_nrst_cavlpnt_xyz i:ssh_stiff.dim i:ssh_stiff.nza "
    "I1:ssh_stiff.rowptr I1:ssh_stiff.colind R1:ssh_stiff.values I1:ssh_stiff.colind_loc I1:ssh_stiff.rowptr_loc "
    "R1:lump2d_south R1:lump2d_north I1:ind_south I1:ind_north i:nn_size I1:nn_num I2:nn_pos R2:hnode R2:hnode_new "
    "R2:zbar_3d_n R2:Z_3d_n R2:Z_3d_n_ib R2:helem R1:bottom_elem_thickness R1:bottom_node_thickness R1:dhe R1:hbar "
    "R1:hbar_old R1:zbar_n_bot R1:zbar_e_bot R1:zbar_n_srf R1:zbar_e_srf R1:coriolis R1:coriolis_node";
const char *kCom = "i:rPEnum S32:rPE S33:rptr I1:rlist i:sPEnum S32:sPE S32:sptr I1:slist i:nreq";
const char *kPartit =
    "i:npes i:mype i:maxPEnum I1:part i:myDim_nod2D i:eDim_nod2D I1:myList_nod2D i:myDim_elem2D i:eDim_elem2D "
    "i:eXDim_elem2D I1:myList_elem2D i:myDim_edge2D i:eDim_edge2D I1:myList_edge2D i:pe_status";
const char *kTracerData =
    "R2:values R3:valuesold R2:valuesAB l:smooth_bh_tra r:gamma0_tra r:gamma1_tra r:gamma2_tra l:i_vert_diff "
    "c:tra_adv_hor c:tra_adv_ver c:tra_adv_lim r:tra_adv_ph r:tra_adv_pv i:ID";
const char *kTracerWork =
    "R2:del_ttf R2:del_ttf_advhoriz R2:del_ttf_advvert R3:dvd_trflx_hor R3:dvd_trflx_ver R2:fct_LO R2:adv_flux_hor "
    "R2:adv_flux_ver R2:fct_ttf_max R2:fct_ttf_min R2:fct_plus R2:fct_minus R3:edge_up_dn_grad I1:nboundary_lay "
    "I2:edge_up_dn_tri";
const char *kDyn =
    "i:opt_visc r:visc_gamma0 r:visc_gamma1 r:visc_gamma2 r:visc_easybsreturn l:use_ivertvisc i:momadv_opt "
    "l:use_freeslip l:use_wsplit r:wsplit_maxcfl l:use_ssh_se_subcycl i:solverinfo.ident i:solverinfo.maxiter "
    "i:solverinfo.restart i:solverinfo.fillin i:solverinfo.lutype r:solverinfo.droptol r:solverinfo.soltol "
    "R1:solverinfo.rr R1:solverinfo.zz R1:solverinfo.pp R1:solverinfo.App R3:work.uvnode_rhs R2:work.u_c "
    "R2:work.v_c R2:work.u_b R2:work.v_b R3:uv R3:uv_rhs R4:uv_rhsAB R3:uvnode R2:w R2:w_e R2:w_i";

struct Spec {
    char kind;   // i r l c
    int rank;    // 0 scalar
    int fixed;   // > 0: static int array of this many elements
    std::string name;
};

std::vector<Spec> parse_specs(const char *s, const std::string &prefix) {
    std::vector<Spec> out;
    std::istringstream in(s);
    std::string tok;
    while (in >> tok) {
        const size_t c = tok.find(':');
        const std::string t = tok.substr(0, c), name = prefix + tok.substr(c + 1);
        Spec sp{'i', 0, 0, name};
        if (t == "i" || t == "r" || t == "l" || t == "c") sp.kind = t[0];
        else if (t[0] == 'I' || t[0] == 'R') {
            sp.kind = t[0] == 'I' ? 'i' : 'r';
            sp.rank = t[1] - '0';
        } else {  // S<n>
            sp.kind = 'i';
            sp.rank = 1;
            sp.fixed = std::stoi(t.substr(1));
        }
        out.push_back(sp);
    }
    return out;
}

size_t elem_size(char k) { return k == 'r' ? 8 : k == 'c' ? 20 : 4; }

struct Reader {
    const std::vector<char> &b;
    size_t pos = 0;
    bool ok = true;
    explicit Reader(const std::vector<char> &buf) : b(buf) {}
    bool take(void *dst, size_t n) {
        if (!ok || pos + n > b.size()) return ok = false;
        std::memcpy(dst, b.data() + pos, n);
        pos += n;
        return true;
    }
};

bool read_item(Reader &r, const Spec &sp, RsField &f) {
    f.kind = sp.kind;
    f.array = sp.rank > 0;
    f.dims.clear();
    size_t n = 1;
    for (int d = 0; d < sp.rank; ++d) {
        int32_t e = 0;
        if (!r.take(&e, 4) || e < 0) return false;
        f.dims.push_back(e);
        n *= (size_t)e;
    }
    f.data.resize(n * elem_size(sp.kind));
    return r.take(f.data.data(), f.data.size());
}

void put(std::vector<char> &o, const void *p, size_t n) {
    const char *c = static_cast<const char *>(p);
    o.insert(o.end(), c, c + n);
}

void write_item(std::vector<char> &o, const Spec &sp, const RsField *f) {
    if (f && f->kind == sp.kind && (int)f->dims.size() == sp.rank) {
        for (int e : f->dims) {
            const int32_t v = e;
            put(o, &v, 4);
        }
        put(o, f->data.data(), f->data.size());
        return;
    }
    if (sp.fixed > 0) {  // a fixed array is always written whole
        const int32_t v = sp.fixed;
        put(o, &v, 4);
        o.insert(o.end(), (size_t)sp.fixed * 4, 0);
    } else if (sp.rank > 0) {  // not allocated: zero extents
        o.insert(o.end(), (size_t)sp.rank * 4, 0);
    } else if (sp.kind == 'c') {
        o.insert(o.end(), 20, ' ');
    } else {
        o.insert(o.end(), elem_size(sp.kind), 0);
    }
}

// Payload of one unformatted sequential record (sub-records joined).
int read_record(const std::string &path, std::vector<char> &payload, std::string &err) {
    FILE *fp = std::fopen(path.c_str(), "rb");
    if (!fp) {
        err = "cannot open " + path + ": " + std::strerror(errno);
        return -1;
    }
    std::vector<char> all;
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, fp)) > 0) all.insert(all.end(), buf, buf + n);
    std::fclose(fp);
    payload.clear();
    size_t pos = 0;
    for (;;) {
        int32_t head = 0, tail = 0;
        if (pos + 4 > all.size()) {
            err = path + ": truncated record marker";
            return -1;
        }
        std::memcpy(&head, all.data() + pos, 4);
        pos += 4;
        const size_t len = (size_t)(head < 0 ? -(int64_t)head : head);
        if (pos + len + 4 > all.size()) {
            err = path + ": record longer than the file";
            return -1;
        }
        payload.insert(payload.end(), all.data() + pos, all.data() + pos + len);
        pos += len;
        std::memcpy(&tail, all.data() + pos, 4);
        pos += 4;
        if (head >= 0) break;  // the last (or only) sub-record has a positive head marker
    }
    return 0;
}

int write_record(const std::string &path, const std::vector<char> &payload, std::string &err) {
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) {
        err = "cannot create " + path + ": " + std::strerror(errno);
        return -1;
    }
    const size_t kMax = 2147483639;  // largest sub-record of the 4-byte convention
    size_t pos = 0;
    bool first = true;
    do {
        const size_t len = std::min(kMax, payload.size() - pos);
        const bool last = pos + len == payload.size();
        const int32_t head = last ? (int32_t)len : -(int32_t)len, tail = first ? (int32_t)len : -(int32_t)len;
        std::fwrite(&head, 4, 1, fp);
        if (len) std::fwrite(payload.data() + pos, 1, len, fp);
        std::fwrite(&tail, 4, 1, fp);
        pos += len;
        first = false;
    } while (pos < payload.size());
    const bool bad = std::ferror(fp) != 0;
    std::fclose(fp);
    if (bad) {
        err = "write error on " + path;
        return -1;
    }
    return 0;
}

int walk(const std::vector<Spec> &specs, Reader &r, RestartFiles &out, int file) {
    for (const Spec &sp : specs) {
        RsField f;
        if (!read_item(r, sp, f)) return -1;
        out.order[file].push_back(sp.name);
        out.f[sp.name] = std::move(f);
    }
    return 0;
}

std::vector<Spec> partit_specs() {
    std::vector<Spec> s;
    for (const char *c : {"com_nod2D.", "com_elem2D.", "com_elem2D_full."}) {
        auto v = parse_specs(kCom, c);
        s.insert(s.end(), v.begin(), v.end());
    }
    auto v = parse_specs(kPartit, "");
    s.insert(s.end(), v.begin(), v.end());
    return s;
}

std::vector<Spec> tracer_specs(int ntr) {
    std::vector<Spec> s = parse_specs("i:num_tracers", "");
    for (int n = 1; n <= ntr; ++n) {
        auto v = parse_specs(kTracerData, "tracer." + std::to_string(n) + ".");
        s.insert(s.end(), v.begin(), v.end());
    }
    auto w = parse_specs(kTracerWork, "work.");
    s.insert(s.end(), w.begin(), w.end());
    return s;
}

bool make_dir(const std::string &p) { return ::mkdir(p.c_str(), 0755) == 0 || errno == EEXIST; }

}  // namespace

std::vector<int> RestartFiles::ints(const std::string &name) const {
    const RsField *x = get(name);
    std::vector<int> v;
    if (!x || x->kind != 'i') return v;
    v.resize(x->data.size() / 4);
    std::memcpy(v.data(), x->data.data(), x->data.size());
    return v;
}
std::vector<double> RestartFiles::reals(const std::string &name) const {
    const RsField *x = get(name);
    std::vector<double> v;
    if (!x || x->kind != 'r') return v;
    v.resize(x->data.size() / 8);
    std::memcpy(v.data(), x->data.data(), x->data.size());
    return v;
}
int RestartFiles::scalar_int(const std::string &name, int dflt) const {
    const std::vector<int> v = ints(name);
    return v.empty() ? dflt : v[0];
}
double RestartFiles::scalar_real(const std::string &name, double dflt) const {
    const std::vector<double> v = reals(name);
    return v.empty() ? dflt : v[0];
}
bool RestartFiles::logical(const std::string &name, bool dflt) const {
    const RsField *x = get(name);
    if (!x || x->kind != 'l' || x->data.size() < 4) return dflt;
    int32_t v;
    std::memcpy(&v, x->data.data(), 4);
    return v != 0;
}
std::string RestartFiles::chars(const std::string &name) const {
    const RsField *x = get(name);
    if (!x || x->kind != 'c') return "";
    std::string s(x->data.begin(), x->data.end());
    while (!s.empty() && (s.back() == ' ' || s.back() == '\0')) s.pop_back();
    return s;
}
void RestartFiles::set_reals(const std::string &name, const std::vector<double> &v, const std::vector<int> &dims) {
    RsField &x = f[name];
    x.kind = 'r';
    x.array = true;
    x.dims = dims;
    x.data.resize(v.size() * 8);
    if (!v.empty()) std::memcpy(x.data.data(), v.data(), x.data.size());
}
void RestartFiles::set_ints(const std::string &name, const std::vector<int> &v, const std::vector<int> &dims) {
    RsField &x = f[name];
    x.kind = 'i';
    x.array = true;
    x.dims = dims;
    x.data.resize(v.size() * 4);
    if (!v.empty()) std::memcpy(x.data.data(), v.data(), x.data.size());
}
void RestartFiles::set_scalar_int(const std::string &name, int v) {
    RsField &x = f[name];
    x = RsField{};
    x.kind = 'i';
    x.data.resize(4);
    std::memcpy(x.data.data(), &v, 4);
}
void RestartFiles::set_scalar_real(const std::string &name, double v) {
    RsField &x = f[name];
    x = RsField{};
    x.kind = 'r';
    x.data.resize(8);
    std::memcpy(x.data.data(), &v, 8);
}
void RestartFiles::set_logical(const std::string &name, bool v) {
    RsField &x = f[name];
    x = RsField{};
    x.kind = 'l';
    const int32_t i = v ? 1 : 0;
    x.data.resize(4);
    std::memcpy(x.data.data(), &i, 4);
}
void RestartFiles::set_chars(const std::string &name, const std::string &v) {
    RsField &x = f[name];
    x = RsField{};
    x.kind = 'c';
    x.data.assign(20, ' ');
    std::memcpy(x.data.data(), v.data(), std::min<size_t>(20, v.size()));
}

int read_restart(const std::string &npdir, bool tracers, RestartFiles &out, std::string &err) {
    out = RestartFiles{};
    std::vector<char> p;
    // t_mesh
    if (read_record(npdir + "/t_mesh.0", p, err)) return -1;
    {
        Reader r(p);
        if (walk(parse_specs(kMesh, "mesh."), r, out, 0) || r.pos != p.size()) {
            err = npdir + "/t_mesh.0: payload does not follow the mesh type's layout (FP64 restart expected)";
            return -1;
        }
    }
    if (read_record(npdir + "/t_partit.0", p, err)) return -1;
    {
        Reader r(p);
        if (walk(partit_specs(), r, out, 1) || r.pos != p.size()) {
            err = npdir + "/t_partit.0: payload does not follow the partition type's layout";
            return -1;
        }
    }
    if (tracers) {
        if (read_record(npdir + "/t_tracer.0", p, err)) return -1;
        Reader r(p);
        int32_t ntr = 0;
        if (p.size() < 4) {
            err = npdir + "/t_tracer.0: empty";
            return -1;
        }
        std::memcpy(&ntr, p.data(), 4);
        if (ntr < 0 || ntr > 4096 || walk(tracer_specs(ntr), r, out, 2) || r.pos != p.size()) {
            err = npdir + "/t_tracer.0: payload does not follow the tracer type's layout";
            return -1;
        }
        out.has_tracers = true;
    }
    if (read_record(npdir + "/t_dynamics.0", p, err)) return -1;
    {
        Reader r(p);
        if (walk(parse_specs(kDyn, "dyn."), r, out, 3)) {
            err = npdir + "/t_dynamics.0: payload does not follow the dynamics type's layout";
            return -1;
        }
        out.dyn_tail.assign(p.begin() + (long)r.pos, p.end());  // cfl_z and the optional parts
    }
    return 0;
}

int write_restart(const std::string &npdir, const std::string &meta, const RestartFiles &in, int ntracers,
                  std::string &err) {
    const size_t slash = npdir.find_last_of('/');
    if (slash != std::string::npos && !make_dir(npdir.substr(0, slash))) {
        err = "cannot create " + npdir.substr(0, slash);
        return -1;
    }
    if (!make_dir(npdir)) {
        err = "cannot create " + npdir;
        return -1;
    }
    const std::vector<Spec> specs[4] = {parse_specs(kMesh, "mesh."), partit_specs(), tracer_specs(ntracers),
                                        parse_specs(kDyn, "dyn.")};
    const char *names[4] = {"t_mesh", "t_partit", "t_tracer", "t_dynamics"};
    for (int k = 0; k < 4; ++k) {
        std::vector<char> o;
        for (const Spec &sp : specs[k]) write_item(o, sp, in.get(sp.name));
        if (k == 3) {
            if (!in.dyn_tail.empty()) o.insert(o.end(), in.dyn_tail.begin(), in.dyn_tail.end());
            else o.insert(o.end(), 8, 0);  // cfl_z not allocated; no optional parts
        }
        if (write_record(npdir + "/" + names[k] + ".0", o, err)) return -1;
    }
    // meta.time as write_all_bin_restarts writes it for (cstep, ctime, cyear) = (0, 0, 0)
    FILE *fp = std::fopen(meta.c_str(), "w");
    if (!fp) {
        err = "cannot create " + meta;
        return -1;
    }
    std::fprintf(fp, "0\n0\n! year: 0\n!   t_mesh\n!   t_partit\n!   t_tracer\n!   t_dynamics\n");
    std::fclose(fp);
    return 0;
}

}  // namespace fesomtra
