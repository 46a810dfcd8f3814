// Derived-type restart files of the reference (lib/io_restart_derivedtype.F90): <dir>/np<npes>/
// t_mesh.<rank>, t_partit.<rank>, t_tracer.<rank>, t_dynamics.<rank> and <dir>/meta.time.
//
// Each file is one Fortran unformatted sequential record (4-byte length before and after; a
// record longer than 2^31-1 bytes is split into sub-records with negative length markers) whose
// payload is what the type's write procedure emits, item after item, with no markers between
// them: scalars at their kind's size, and every allocatable array as its extents (int32 each)
// followed by its elements in Fortran order, or as zero extents when it is not allocated
// (lib/MOD_WRITE_BINARY_ARRAYS.F90).  The item sequences follow write_t_mesh
// (lib/MOD_MESH.F90:185-279), WRITE_T_PARTIT / WRITE_T_COM_STRUCT (lib/MOD_PARTIT.F90:114-185),
// WRITE_T_TRACER / _DATA / _WORK (lib/MOD_TRACER.F90:104-224) and WRITE_T_DYN with its solver
// and work parts (lib/MOD_DYN.F90).  Reals are 8 bytes (the reference's dp build), integers
// and logicals 4.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace fesomtra {

// One item of a restart payload: kind, array extents (empty for a scalar) and the raw bytes.
struct RsField {
    char kind = 'i';          // 'i' int32, 'r' real64, 'l' logical32, 'c' character(20)
    std::vector<int> dims;    // array extents; scalars: none
    bool array = false;
    std::vector<char> data;   // elements, Fortran order
};

// The four payloads as ordered items (names like "mesh.nod2D", "tracer.2.values"); the part of
// t_dynamics after w_i is kept as raw bytes ("dyn.tail").
struct RestartFiles {
    std::vector<std::string> order[4];  // mesh, partit, tracer, dynamics
    std::map<std::string, RsField> f;
    std::vector<char> dyn_tail;
    bool has_tracers = false;

    const RsField *get(const std::string &name) const {
        auto it = f.find(name);
        return it == f.end() ? nullptr : &it->second;
    }
    std::vector<int> ints(const std::string &name) const;
    std::vector<double> reals(const std::string &name) const;
    int scalar_int(const std::string &name, int dflt = 0) const;
    double scalar_real(const std::string &name, double dflt = 0.0) const;
    bool logical(const std::string &name, bool dflt = false) const;
    std::string chars(const std::string &name) const;  // trailing blanks removed
    void set_reals(const std::string &name, const std::vector<double> &v, const std::vector<int> &dims);
    void set_ints(const std::string &name, const std::vector<int> &v, const std::vector<int> &dims);
    void set_scalar_int(const std::string &name, int v);
    void set_scalar_real(const std::string &name, double v);
    void set_logical(const std::string &name, bool v);
    void set_chars(const std::string &name, const std::string &v);
};

// Reads <npdir>/t_mesh.0, t_partit.0, t_dynamics.0 and, with tracers, t_tracer.0.
int read_restart(const std::string &npdir, bool tracers, RestartFiles &out, std::string &err);
// Writes the four files into <npdir> (created) and <meta>, items in the write procedures'
// order: fields present in `in` are written from it, the others as not allocated (arrays) or
// zero (scalars).  `ntracers` sets the number of tracer blocks of t_tracer.
int write_restart(const std::string &npdir, const std::string &meta, const RestartFiles &in, int ntracers,
                  std::string &err);

}  // namespace fesomtra
