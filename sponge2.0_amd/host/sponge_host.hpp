// sponge_host.hpp -- the C++ host side of the nonbonded path: SPONGE's
// controller and module interfaces (the method names, argument meaning and
// mdin keys of the reference's CONTROLLER, MD_INFORMATION, NEIGHBOR_LIST,
// LENNARD_JONES_INFORMATION, Particle_Mesh, NON_BOND_14, BOND,
// MIDDLE_Langevin_INFORMATION and SETTLE) with their bodies calling
// libsponge_nb through its C ABI (include/sponge_nb.h).  One snb_context
// holds every module's device state; device buffers of the MD state (crd,
// vel, frc, ...) are the host's, as MD_INFORMATION's are in the reference.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/sponge_nb.h"

typedef snb_vector VECTOR;
typedef snb_ltmatrix3 LTMatrix3;

// CONTROLLER (control.h, control.cpp, utils/control/print.hpp): the mdin
// command table and the mdout columns.
struct CONTROLLER
{
    std::map<std::string, std::string> commands;
    std::vector<std::string> outputs_key;
    std::map<std::string, std::string> outputs_format, outputs_content;
    FILE* mdout = nullptr;
    float printf_sum = 0.0f;

    // mdin text format (control.cpp:365-431): first line = md_name; then
    // "key = value" lines, "prefix { ... }" blocks giving prefix_key, '#'
    // comments.  argv: [-mdin file] and "-key value" overrides.
    void Initial(int argc, char** argv);
    void Read_Mdin(const char* path);
    bool Command_Exist(const char* key) const;
    bool Command_Exist(const char* prefix, const char* key) const;
    const char* Command(const char* key) const;
    const char* Command(const char* prefix, const char* key) const;
    float Get_Float(const char* key, float dflt) const;
    int Get_Int(const char* key, int dflt) const;
    bool Get_Bool(const char* key, bool dflt) const;
    std::string In_File(const char* key) const;  // "<key>" or default_in_file_prefix_<name>.txt
    void Step_Print_Initial(const char* head, const char* format);
    void Step_Print(const char* head, float value, bool add_to_total = false);
    void Step_Print(const char* head, double value, bool add_to_total = false);
    void Step_Print(const char* head, int value);
    void Print_First_Line_To_Mdout();
    void Print_To_Screen_And_Mdout();
    [[noreturn]] void Throw_SPONGE_Error(const char* where, const char* reason) const;
    void Check(int snb_rc, const char* where) const;  // snb status -> Throw_SPONGE_Error
};

// Native input files (xponge/load/native/*.hpp formats)
struct NativeSystem
{
    int atom_numbers = 0;
    std::vector<float> crd, vel, mass, charge;  // crd/vel: 3 per atom
    VECTOR box_length{0, 0, 0}, box_angle{90, 90, 90};
    int lj_types = 0;
    std::vector<float> lj_a, lj_b;  // x12 / x6 applied (lj.hpp:47, :57)
    std::vector<int> lj_type;
    std::vector<int> ex_start, ex_num, ex_list;
    std::vector<int> nb14_a, nb14_b;
    std::vector<float> nb14_A, nb14_B, nb14_cf;
    std::vector<int> bond_a, bond_b;
    std::vector<float> bond_k, bond_r0;
    bool has_velocity = false;
};
void Load_Native_System(CONTROLLER* controller, NativeSystem* sys);

// LTMatrix3 cell / rcell of a box (MD_core/pbc.hpp:66-101)
void Box_To_Cell(VECTOR box, VECTOR angle, LTMatrix3* cell, LTMatrix3* rcell);

// MD_INFORMATION (MD_core/MD_core.h): the MD state in device memory
struct MD_INFORMATION
{
    enum Mode { NVE = 0, NVT = 1 };
    int atom_numbers = 0;
    int mode = NVE;
    int steps = 0, step_limit = 1000, write_mdout_interval = 1000, write_frc_interval = 1000;
    float dt_in_ps = 0.001f, dt = 0.0f;  // dt in the internal unit (x 20.455)
    float target_temperature = 300.0f;
    int freedom = 0;
    VECTOR box_length;
    LTMatrix3 cell, rcell;
    float volume = 0.0f;
    VECTOR *crd = nullptr, *vel = nullptr, *frc = nullptr;
    float *d_mass = nullptr, *d_inverse_mass = nullptr, *d_charge = nullptr;
    float *d_atom_energy = nullptr, *d_potential = nullptr;
    LTMatrix3 *d_atom_virial = nullptr, *d_virial = nullptr;
    std::vector<float> h_mass;
    FILE* frc_traj = nullptr;
    void Initial(CONTROLLER* controller, const NativeSystem& sys, hipStream_t stream);
    double Kinetic_Energy(hipStream_t stream) const;
    void Step_Print(CONTROLLER* controller, hipStream_t stream);
    void Append_Frc_Traj_File(hipStream_t stream);
    void Clear();
};

struct NEIGHBOR_LIST
{
    float cutoff = 10.0f, skin = 2.0f;
    int refresh_interval = 0;
    snb_atom_group* d_nl = nullptr;  // the reference-layout list (public d_nl)
    void Initial(CONTROLLER* controller, snb_context* ctx, float cutoff, float skin,
                 LTMatrix3 cell, LTMatrix3 rcell);
    void Update(snb_context* ctx, int atom_numbers, const VECTOR* crd, LTMatrix3 cell,
                LTMatrix3 rcell, int step, int forced);
};

struct LENNARD_JONES_INFORMATION
{
    bool is_initialized = false;
    void Initial(CONTROLLER* controller, snb_context* ctx, const NativeSystem& sys,
                 float cutoff);
    void LJ_PME_Direct_Force_With_Atom_Energy_And_Virial(
        snb_context* ctx, int atom_numbers, const VECTOR* crd, const float* charge, VECTOR* frc,
        LTMatrix3 cell, LTMatrix3 rcell, const snb_atom_group* nl, float pme_beta,
        int need_atom_energy, float* atom_energy, int need_virial, LTMatrix3* atom_virial);
    void Long_Range_Correction(snb_context* ctx, int need_pressure, LTMatrix3* d_virial,
                               int need_potential, float* d_potential, float volume);
    void Step_Print(CONTROLLER* controller, snb_context* ctx);
};

struct Particle_Mesh
{
    bool is_initialized = false, print_detail = false;
    float beta = 0.0f;
    void Initial(CONTROLLER* controller, snb_context* ctx, LTMatrix3 cell, LTMatrix3 rcell,
                 VECTOR box_length, float cutoff);
    void PME_Excluded_Force_With_Atom_Energy(snb_context* ctx, int atom_numbers,
                                             const VECTOR* crd, LTMatrix3 cell, LTMatrix3 rcell,
                                             const float* charge, VECTOR* frc, int need_energy,
                                             float* atom_energy, int need_virial,
                                             LTMatrix3* atom_virial);
    void PME_Reciprocal_Force_With_Energy_And_Virial(snb_context* ctx, const VECTOR* crd,
                                                     LTMatrix3 cell, LTMatrix3 rcell,
                                                     const float* charge, VECTOR* frc,
                                                     int need_virial, int need_energy,
                                                     LTMatrix3* d_virial, float* d_potential,
                                                     int step);
    void Step_Print(CONTROLLER* controller, snb_context* ctx);
};

struct NON_BOND_14
{
    bool is_initialized = false;
    void Initial(CONTROLLER* controller, snb_context* ctx, const NativeSystem& sys);
    void Non_Bond_14_LJ_CF_Force_With_Atom_Energy_And_Virial(
        snb_context* ctx, const VECTOR* crd, LTMatrix3 cell, LTMatrix3 rcell, const float* charge,
        VECTOR* frc, int need_energy, float* atom_energy, int need_virial,
        LTMatrix3* atom_virial);
    void Step_Print(CONTROLLER* controller, snb_context* ctx);
};

struct BOND
{
    bool is_initialized = false;
    void Initial(CONTROLLER* controller, snb_context* ctx, const NativeSystem& sys);
    void Bond_Force_With_Atom_Energy_And_Virial(snb_context* ctx, const VECTOR* crd,
                                                LTMatrix3 cell, LTMatrix3 rcell, VECTOR* frc,
                                                int need_energy, float* atom_energy,
                                                int need_virial, LTMatrix3* atom_virial);
    void Step_Print(CONTROLLER* controller, snb_context* ctx);
};

// SETTLE (constrain/settle.cpp) on the rigid triangles and rigid pairs found
// among the constraint pairs (bonds with an atom lighter than 3.3,
// constrain.cpp:78-117)
struct SETTLE
{
    bool is_initialized = false;
    int triangle_numbers = 0, pair_numbers = 0, constrain_pair_numbers = 0;
    void Initial(CONTROLLER* controller, snb_context* ctx, const NativeSystem& sys);
};

struct MIDDLE_Langevin_INFORMATION
{
    bool is_initialized = false;
    float exp_gamma = 1.0f;
    void Initial(CONTROLLER* controller, snb_context* ctx, const NativeSystem& sys,
                 float target_temperature);
};
