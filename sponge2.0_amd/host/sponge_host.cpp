// sponge_host.cpp -- bodies of the controller and module interfaces declared
// in sponge_host.hpp: mdin parsing and mdout printing in the reference's
// formats, the native input readers, and each module method as calls into
// libsponge_nb.
#include "sponge_host.hpp"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <fstream>

static const double kKB = 0.00198716;         // kcal/mol/K, common.h:122
static const float kTIME_UNIT = 20.455f;      // 1 ps in the internal unit, common.h:136

// ------------------------------------------------------------------ CONTROLLER
static std::string trim(const std::string& s)
{
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// The mdin text format as control.cpp:365-431 and its read_one_line
// (:238-283) read it: after the first line (md_name) the file is a stream of
// items ended by '\n', ',', '{', '}' or '\r'; an item ended by '{' names a
// block whose keys become <block>_<key> until '}'; "key = value" items set a
// command (first word of the key, value trimmed); '#' starts a comment.
void CONTROLLER::Read_Mdin(const char* path)
{
    FILE* f = fopen(path, "r");
    if (!f) Throw_SPONGE_Error("CONTROLLER::Initial", "cannot open the mdin file");
    std::string text;
    char buf[4096];
    size_t got;
    while ((got = fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
    fclose(f);
    size_t pos = text.find('\n');
    commands["md_name"] = trim(text.substr(0, pos));
    pos = pos == std::string::npos ? text.size() : pos + 1;
    auto delim = [](char c) { return c == '\n' || c == ',' || c == '{' || c == '}' || c == '\r'; };
    std::string prefix;
    while (pos < text.size())
    {
        // one item and the delimiters (and blanks) after it
        while (pos < text.size() && (text[pos] == ' ' || text[pos] == '\t')) pos++;
        std::string item, ender;
        while (pos < text.size() && !delim(text[pos])) item += text[pos++];
        while (pos < text.size())
        {
            const char c = text[pos];
            if (c == ' ' || c == '\t')
                pos++;
            else if (delim(c))
            {
                ender += c;
                pos++;
            }
            else
                break;
        }
        if (!item.empty() && item[0] == '#')
        {
            if (item.size() > 1 && item[1] == '#' && ender.find('{') != std::string::npos)
                prefix = trim(item).substr(0, trim(item).find_first_of(" \t"));
            if (ender.find('\n') == std::string::npos)
            {
                // the rest of the comment line
                while (pos < text.size() && text[pos] != '\n') pos++;
            }
        }
        else if (ender.find('{') != std::string::npos)
        {
            std::string t = trim(item);
            prefix = t.substr(0, t.find_first_of(" \t"));
        }
        else if ((prefix.size() < 2 || prefix[0] != '#' || prefix[1] != '#'))
        {
            // Get_Command (control.cpp:207-236)
            size_t eq = item.find('=');
            if (eq != std::string::npos)
            {
                std::string key = trim(item.substr(0, eq));
                key = key.substr(0, key.find_first_of(" \t"));
                std::string value = trim(item.substr(eq + 1));
                size_t eq2 = value.find('=');
                if (eq2 != std::string::npos) value = trim(value.substr(0, eq2));
                size_t hash = value.find('#');  // a trailing comment is not part of the value
                if (hash != std::string::npos) value = trim(value.substr(0, hash));
                if (!key.empty() && !value.empty())
                    commands[prefix.empty() ? key : prefix + "_" + key] = value;
            }
        }
        if (ender.find('}') != std::string::npos) prefix.clear();
    }
}

void CONTROLLER::Initial(int argc, char** argv)
{
    std::string mdin = "mdin.txt";
    for (int i = 1; i + 1 < argc; i += 2)
        if (strcmp(argv[i], "-mdin") == 0) mdin = argv[i + 1];
    Read_Mdin(mdin.c_str());
    for (int i = 1; i + 1 < argc; i += 2)
        if (argv[i][0] == '-' && strcmp(argv[i], "-mdin") != 0) commands[argv[i] + 1] = argv[i + 1];
    if (Command_Exist("dump_commands"))
    {
        // the parsed command table, one "key = value" per line, then stop
        for (const auto& kv : commands) printf("%s = %s\n", kv.first.c_str(), kv.second.c_str());
        exit(0);
    }
    const char* out = Command_Exist("mdout") ? Command("mdout") : "mdout.txt";
    mdout = fopen(out, "w");
    if (!mdout) Throw_SPONGE_Error("CONTROLLER::Initial", "cannot open the mdout file");
}

bool CONTROLLER::Command_Exist(const char* key) const { return commands.count(key) > 0; }

bool CONTROLLER::Command_Exist(const char* prefix, const char* key) const
{
    return commands.count(std::string(prefix) + "_" + key) > 0;
}

const char* CONTROLLER::Command(const char* key) const
{
    auto it = commands.find(key);
    if (it == commands.end()) Throw_SPONGE_Error("CONTROLLER::Command", key);
    return it->second.c_str();
}

const char* CONTROLLER::Command(const char* prefix, const char* key) const
{
    return Command((std::string(prefix) + "_" + key).c_str());
}

float CONTROLLER::Get_Float(const char* key, float dflt) const
{
    return Command_Exist(key) ? (float)atof(Command(key)) : dflt;
}

int CONTROLLER::Get_Int(const char* key, int dflt) const
{
    return Command_Exist(key) ? atoi(Command(key)) : dflt;
}

bool CONTROLLER::Get_Bool(const char* key, bool dflt) const
{
    if (!Command_Exist(key)) return dflt;
    std::string v = Command(key);
    std::transform(v.begin(), v.end(), v.begin(), ::tolower);
    return v == "1" || v == "true" || v == "yes" || v == "on";
}

std::string CONTROLLER::In_File(const char* key) const
{
    std::string k = std::string(key) + "_in_file";
    if (Command_Exist(k.c_str())) return Command(k.c_str());
    // Command_Exist for *_in_file keys (control.cpp:49-75): <prefix>_<name>.txt
    if (Command_Exist("default_in_file_prefix"))
    {
        std::string p = std::string(Command("default_in_file_prefix")) + "_" + key + ".txt";
        FILE* f = fopen(p.c_str(), "r");
        if (f)
        {
            fclose(f);
            return p;
        }
    }
    return std::string();
}

// utils/control/print.hpp:71-196: columns registered in order, each printed
// with its own format into a %15s field; "****" until a value is set
void CONTROLLER::Step_Print_Initial(const char* head, const char* format)
{
    outputs_format[head] = format;
    outputs_content[head] = "****";
    outputs_key.push_back(head);
}

void CONTROLLER::Step_Print(const char* head, float value, bool add_to_total)
{
    if (!outputs_content.count(head)) return;
    char t[256];
    snprintf(t, sizeof(t), outputs_format[head].c_str(), value);
    outputs_content[head] = t;
    if (add_to_total) printf_sum += value;
}

void CONTROLLER::Step_Print(const char* head, double value, bool add_to_total)
{
    if (!outputs_content.count(head)) return;
    char t[256];
    snprintf(t, sizeof(t), outputs_format[head].c_str(), value);
    outputs_content[head] = t;
    if (add_to_total) printf_sum += (float)value;
}

void CONTROLLER::Step_Print(const char* head, int value)
{
    if (!outputs_content.count(head)) return;
    char t[256];
    snprintf(t, sizeof(t), outputs_format[head].c_str(), value);
    outputs_content[head] = t;
}

void CONTROLLER::Print_First_Line_To_Mdout()
{
    for (size_t i = 0; i < outputs_key.size(); i++)
        fprintf(mdout, "%15s%s", outputs_key[i].c_str(), i + 1 == outputs_key.size() ? "\n" : " ");
    printf("---------------------------------------------------------------------------------"
           "---------------------------\n");
}

void CONTROLLER::Print_To_Screen_And_Mdout()
{
    int count = 0;
    for (size_t i = 0; i < outputs_key.size(); i++)
    {
        const std::string& k = outputs_key[i];
        count++;
        printf("%15s = %15s, ", k.c_str(), outputs_content[k].c_str());
        fprintf(mdout, "%15s%s", outputs_content[k].c_str(),
                i + 1 == outputs_key.size() ? "\n" : " ");
        outputs_content[k] = "****";
        if (count % 3 == 0) printf("\n");
    }
    if (count % 3 != 0) printf("\n");
    printf("---------------------------------------------------------------------------------"
           "---------------------------\n");
    printf_sum = 0.0f;
    fflush(mdout);
}

void CONTROLLER::Throw_SPONGE_Error(const char* where, const char* reason) const
{
    fprintf(stderr, "\nSPONGE ERROR in %s\n\t%s\n", where, reason);
    exit(1);
}

void CONTROLLER::Check(int rc, const char* where) const
{
    if (rc != SNB_OK) Throw_SPONGE_Error(where, snb_status_string(rc));
}

// ------------------------------------------------------------- native inputs
static std::vector<std::string> tokens(CONTROLLER* c, const std::string& path)
{
    std::ifstream in(path);
    if (!in) c->Throw_SPONGE_Error("Load_Native_System", ("cannot open " + path).c_str());
    std::vector<std::string> t;
    std::string w;
    while (in >> w) t.push_back(w);
    return t;
}

// coordinate / velocity file (md_core.hpp:90-158): "N [t0]", xyz per atom,
// then (coordinate only) the box lengths and angles
static void read_xyz(CONTROLLER* c, const std::string& path, int n_expect, std::vector<float>* xyz,
                     VECTOR* box, VECTOR* angle)
{
    std::ifstream in(path);
    if (!in) c->Throw_SPONGE_Error("Load_Native_System", ("cannot open " + path).c_str());
    std::string first;
    std::getline(in, first);
    int n = atoi(first.c_str());
    if (n_expect >= 0 && n != n_expect)
        c->Throw_SPONGE_Error("Load_Native_System", ("atom count differs in " + path).c_str());
    xyz->resize(3 * (size_t)n);
    for (size_t k = 0; k < 3 * (size_t)n; k++) in >> (*xyz)[k];
    if (box)
    {
        in >> box->x >> box->y >> box->z;
        if (!(in >> angle->x >> angle->y >> angle->z)) *angle = {90.0f, 90.0f, 90.0f};
    }
}

void Load_Native_System(CONTROLLER* c, NativeSystem* s)
{
    std::string f = c->In_File("coordinate");
    if (f.empty()) c->Throw_SPONGE_Error("Load_Native_System", "coordinate_in_file is required");
    read_xyz(c, f, -1, &s->crd, &s->box_length, &s->box_angle);
    const int n = s->atom_numbers = (int)(s->crd.size() / 3);
    if (!(f = c->In_File("velocity")).empty())
    {
        read_xyz(c, f, n, &s->vel, nullptr, nullptr);
        s->has_velocity = true;
    }
    else
        s->vel.assign(3 * (size_t)n, 0.0f);
    auto scalars = [&](const char* key, std::vector<float>* out, float dflt) {
        std::string p = c->In_File(key);
        out->assign(n, dflt);
        if (p.empty()) return;
        auto t = tokens(c, p);
        for (int i = 0; i < n && i + 1 < (int)t.size(); i++) (*out)[i] = (float)atof(t[1 + i].c_str());
    };
    scalars("mass", &s->mass, 20.0f);
    scalars("charge", &s->charge, 0.0f);
    // LJ (lj.hpp:8-69): "N T", A lower triangle, B lower triangle, type per atom
    if ((f = c->In_File("LJ")).empty())
        c->Throw_SPONGE_Error("Load_Native_System", "LJ_in_file is required");
    {
        auto t = tokens(c, f);
        const int nt = s->lj_types = atoi(t[1].c_str());
        const size_t np = (size_t)nt * (nt + 1) / 2;
        s->lj_a.resize(np);
        s->lj_b.resize(np);
        s->lj_type.resize(n);
        for (size_t k = 0; k < np; k++)
        {
            s->lj_a[k] = (float)atof(t[2 + k].c_str()) * 12.0f;
            s->lj_b[k] = (float)atof(t[2 + np + k].c_str()) * 6.0f;
        }
        for (int i = 0; i < n; i++) s->lj_type[i] = atoi(t[2 + 2 * np + i].c_str());
    }
    // exclude (md_core.hpp:230-284): "N total", per atom "count j..."
    s->ex_start.assign(n, 0);
    s->ex_num.assign(n, 0);
    if (!(f = c->In_File("exclude")).empty())
    {
        auto t = tokens(c, f);
        size_t pos = 2;
        for (int i = 0; i < n; i++)
        {
            const int cnt = atoi(t[pos++].c_str());
            s->ex_start[i] = (int)s->ex_list.size();
            s->ex_num[i] = cnt;
            for (int k = 0; k < cnt; k++) s->ex_list.push_back(atoi(t[pos++].c_str()));
        }
    }
    // nb14 (nb14.hpp:8-95): extra pairs "a b A B cf" first, then "a b lj_scale cf"
    if (!(f = c->In_File("nb14_extra")).empty())
    {
        auto t = tokens(c, f);
        const int m = atoi(t[0].c_str());
        for (int k = 0; k < m; k++)
        {
            s->nb14_a.push_back(atoi(t[1 + 5 * k].c_str()));
            s->nb14_b.push_back(atoi(t[2 + 5 * k].c_str()));
            s->nb14_A.push_back((float)atof(t[3 + 5 * k].c_str()) * 12.0f);
            s->nb14_B.push_back((float)atof(t[4 + 5 * k].c_str()) * 6.0f);
            s->nb14_cf.push_back((float)atof(t[5 + 5 * k].c_str()));
        }
    }
    if (!(f = c->In_File("nb14")).empty())
    {
        auto t = tokens(c, f);
        const int m = atoi(t[0].c_str());
        for (int k = 0; k < m; k++)
        {
            const int a = atoi(t[1 + 4 * k].c_str()), b = atoi(t[2 + 4 * k].c_str());
            const float sc = (float)atof(t[3 + 4 * k].c_str());
            const int ta = s->lj_type[a], tb = s->lj_type[b];
            const int hi = std::max(ta, tb), lo = std::min(ta, tb);
            const size_t p = (size_t)hi * (hi + 1) / 2 + lo;
            s->nb14_a.push_back(a);
            s->nb14_b.push_back(b);
            s->nb14_A.push_back(sc * s->lj_a[p]);
            s->nb14_B.push_back(sc * s->lj_b[p]);
            s->nb14_cf.push_back((float)atof(t[4 + 4 * k].c_str()));
        }
    }
    // bond (xponge/load/native/bond.hpp:8-40): "M", per bond "a b k r0"
    if (!(f = c->In_File("bond")).empty())
    {
        auto t = tokens(c, f);
        const int m = atoi(t[0].c_str());
        for (int k = 0; k < m; k++)
        {
            s->bond_a.push_back(atoi(t[1 + 4 * k].c_str()));
            s->bond_b.push_back(atoi(t[2 + 4 * k].c_str()));
            s->bond_k.push_back((float)atof(t[3 + 4 * k].c_str()));
            s->bond_r0.push_back((float)atof(t[4 + 4 * k].c_str()));
        }
    }
}

void Box_To_Cell(VECTOR box, VECTOR angle, LTMatrix3* cell, LTMatrix3* rcell)
{
    const float d2r = 0.0174532925f;
    const float al = angle.x * d2r, be = angle.y * d2r, ga = angle.z * d2r;
    const float cal = cosf(al), cbe = cosf(be), cga = cosf(ga), sga = sinf(ga);
    const float za = sqrtf(1.0f - cal * cal - cbe * cbe - cga * cga + 2.0f * cal * cbe * cga);
    float c[6] = {box.x, box.y * cga, box.y * sga, box.z * cbe, box.z / sga * (cal - cbe * cga),
                  box.z / sga * za};
    float r[6];
    r[0] = 1.0f / c[0];
    r[2] = 1.0f / c[2];
    r[5] = 1.0f / c[5];
    r[1] = -r[0] / tanf(ga);
    r[3] = (cal / tanf(ga) - cbe / sga) / za / box.x;
    r[4] = (cbe / tanf(ga) - cal / sga) / za / box.y;
    for (int k : {1, 3, 4})
    {
        if (fabsf(c[k]) < 1e-3f) c[k] = 0.0f;
        if (fabsf(r[k]) < 1e-3f) r[k] = 0.0f;
    }
    *cell = {c[0], c[1], c[2], c[3], c[4], c[5]};
    *rcell = {r[0], r[1], r[2], r[3], r[4], r[5]};
}

// ------------------------------------------------------------ MD_INFORMATION
template <class T>
static T* device_copy(CONTROLLER* c, const T* h, size_t n)
{
    T* d = nullptr;
    if (hipMalloc((void**)&d, sizeof(T) * (n ? n : 1)) != hipSuccess)
        c->Throw_SPONGE_Error("MD_INFORMATION::Initial", "device allocation failed");
    if (n && h && hipMemcpy(d, h, sizeof(T) * n, hipMemcpyHostToDevice) != hipSuccess)
        c->Throw_SPONGE_Error("MD_INFORMATION::Initial", "copy to the device failed");
    if (n && !h && hipMemset(d, 0, sizeof(T) * n) != hipSuccess)
        c->Throw_SPONGE_Error("MD_INFORMATION::Initial", "device clear failed");
    return d;
}

void MD_INFORMATION::Initial(CONTROLLER* c, const NativeSystem& s, hipStream_t stream)
{
    atom_numbers = s.atom_numbers;
    std::string m = c->Command_Exist("mode") ? c->Command("mode") : "NVE";
    std::transform(m.begin(), m.end(), m.begin(), ::toupper);
    if (m == "NVE")
        mode = NVE;
    else if (m == "NVT")
        mode = NVT;
    else
        c->Throw_SPONGE_Error("MD_INFORMATION::Initial", "mode must be NVE or NVT here");
    dt_in_ps = c->Get_Float("dt", 0.001f);
    dt = dt_in_ps * kTIME_UNIT;
    step_limit = c->Get_Int("step_limit", 1000);
    const int info = c->Get_Int("write_information_interval", 1000);
    write_mdout_interval = c->Get_Int("write_mdout_interval", info);
    write_frc_interval = c->Get_Int("write_trajectory_interval", info);
    target_temperature = c->Get_Float("target_temperature", 300.0f);
    box_length = s.box_length;
    Box_To_Cell(s.box_length, s.box_angle, &cell, &rcell);
    volume = cell.a11 * cell.a22 * cell.a33;
    freedom = 3 * atom_numbers;
    h_mass = s.mass;
    std::vector<float> inv(atom_numbers);
    for (int i = 0; i < atom_numbers; i++) inv[i] = h_mass[i] > 0.0f ? 1.0f / h_mass[i] : 0.0f;
    crd = device_copy(c, (const VECTOR*)s.crd.data(), atom_numbers);
    vel = device_copy(c, (const VECTOR*)s.vel.data(), atom_numbers);
    frc = device_copy<VECTOR>(c, nullptr, atom_numbers);
    d_mass = device_copy(c, h_mass.data(), atom_numbers);
    d_inverse_mass = device_copy(c, inv.data(), atom_numbers);
    d_charge = device_copy(c, s.charge.data(), atom_numbers);
    d_atom_energy = device_copy<float>(c, nullptr, atom_numbers);
    d_atom_virial = device_copy<LTMatrix3>(c, nullptr, atom_numbers);
    d_virial = device_copy<LTMatrix3>(c, nullptr, 1);
    d_potential = device_copy<float>(c, nullptr, 1);
    (void)stream;
    // columns of MD_INFORMATION::system_information::Initial (sys.hpp:618-620)
    c->Step_Print_Initial("step", "%d");
    c->Step_Print_Initial("time", "%.3lf");
    c->Step_Print_Initial("temperature", "%.2f");
    c->Step_Print_Initial("potential", "%.2f");  // main.cpp:89
    if (c->Command_Exist("frc"))
    {
        frc_traj = fopen(c->Command("frc"), "wb");
        if (!frc_traj) c->Throw_SPONGE_Error("MD_INFORMATION::Initial", "cannot open frc");
    }
}

double MD_INFORMATION::Kinetic_Energy(hipStream_t stream) const
{
    std::vector<VECTOR> v(atom_numbers);
    (void)hipMemcpyAsync(v.data(), vel, sizeof(VECTOR) * atom_numbers, hipMemcpyDeviceToHost,
                         stream);
    (void)hipStreamSynchronize(stream);
    double ek = 0.0;
    for (int i = 0; i < atom_numbers; i++)
        ek += 0.5 * h_mass[i] * ((double)v[i].x * v[i].x + (double)v[i].y * v[i].y +
                                 (double)v[i].z * v[i].z);
    return ek;
}

// MD_INFORMATION::Step_Print (MD_core.cpp:911-923): step, time, temperature
// (2 E_k / k_B / freedom, sys.hpp:474)
void MD_INFORMATION::Step_Print(CONTROLLER* c, hipStream_t stream)
{
    c->Step_Print("step", steps);
    c->Step_Print("time", (double)dt_in_ps * steps);
    c->Step_Print("temperature", (float)(Kinetic_Energy(stream) * 2.0 / kKB / freedom));
}

// trajectory_output::Append_Frc_Traj_File (MD_core/output.hpp:148-162):
// float32 N x 3 per frame, no header
void MD_INFORMATION::Append_Frc_Traj_File(hipStream_t stream)
{
    if (!frc_traj) return;
    std::vector<VECTOR> f(atom_numbers);
    (void)hipMemcpyAsync(f.data(), frc, sizeof(VECTOR) * atom_numbers, hipMemcpyDeviceToHost,
                         stream);
    (void)hipStreamSynchronize(stream);
    fwrite(f.data(), sizeof(VECTOR), atom_numbers, frc_traj);
    fflush(frc_traj);
}

void MD_INFORMATION::Clear()
{
    void* bufs[] = {crd, vel, frc, d_mass, d_inverse_mass, d_charge, d_atom_energy,
                    d_atom_virial, d_virial, d_potential};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    if (frc_traj) fclose(frc_traj);
}

// ------------------------------------------------------------- NEIGHBOR_LIST
void NEIGHBOR_LIST::Initial(CONTROLLER* c, snb_context* ctx, float cutoff_, float skin_,
                            LTMatrix3 cell, LTMatrix3 rcell)
{
    cutoff = cutoff_;
    skin = skin_;
    // neighbor_list.cpp:746-790 keys and defaults
    const int max_nb = c->Get_Int("neighbor_list_max_neighbor_numbers", 1200);
    const int max_cell = c->Get_Int("neighbor_list_max_atom_in_grid_numbers", 150);
    refresh_interval = c->Get_Int("neighbor_list_refresh_interval", 0);
    const float permit = c->Get_Float("neighbor_list_skin_permit", 0.5f);
    c->Check(snb_nl_initial(ctx, cutoff, skin, cell, rcell, max_nb, max_cell, refresh_interval,
                            permit),
             "NEIGHBOR_LIST::Initial");
    int* serial = nullptr;
    int width = 0;
    c->Check(snb_nl_device_list(ctx, &d_nl, &serial, &width), "NEIGHBOR_LIST::Initial");
    // nothing of this driver reads the reference-layout list
    c->Check(snb_nl_set_export(ctx, c->Get_Bool("neighbor_list_export", false) ? 1 : 0),
             "NEIGHBOR_LIST::Initial");
}

void NEIGHBOR_LIST::Update(snb_context* ctx, int atom_numbers, const VECTOR* crd, LTMatrix3 cell,
                           LTMatrix3 rcell, int step, int forced)
{
    (void)snb_nl_update(ctx, nullptr, atom_numbers, 0, crd, cell, rcell, step, forced, nullptr,
                        nullptr, nullptr);
}

// ---------------------------------------------------- LENNARD_JONES_INFORMATION
void LENNARD_JONES_INFORMATION::Initial(CONTROLLER* c, snb_context* ctx, const NativeSystem& s,
                                        float cutoff)
{
    c->Check(snb_lj_initial(ctx, s.lj_types, s.lj_a.data(), s.lj_b.data(), s.lj_type.data(),
                            cutoff),
             "LENNARD_JONES_INFORMATION::Initial");
    // Lennard_Jones_force.cpp:275-277
    c->Step_Print_Initial("LJ_short", "%.2f");
    c->Step_Print_Initial("LJ_long", "%.2f");
    c->Step_Print_Initial("LJ", "%.2f");
    is_initialized = true;
}

void LENNARD_JONES_INFORMATION::LJ_PME_Direct_Force_With_Atom_Energy_And_Virial(
    snb_context* ctx, int atom_numbers, const VECTOR* crd, const float* charge, VECTOR* frc,
    LTMatrix3 cell, LTMatrix3 rcell, const snb_atom_group* nl, float pme_beta,
    int need_atom_energy, float* atom_energy, int need_virial, LTMatrix3* atom_virial)
{
    (void)snb_lj_pme_direct_force(ctx, atom_numbers, atom_numbers, 0, 0, crd, charge, frc, cell,
                                  rcell, nl, pme_beta, need_atom_energy, atom_energy, need_virial,
                                  atom_virial, nullptr);
}

void LENNARD_JONES_INFORMATION::Long_Range_Correction(snb_context* ctx, int need_pressure,
                                                      LTMatrix3* d_virial, int need_potential,
                                                      float* d_potential, float volume)
{
    (void)snb_lj_long_range_correction(ctx, need_pressure, d_virial, need_potential, d_potential,
                                       volume);
}

// Lennard_Jones_force.cpp:419-433
void LENNARD_JONES_INFORMATION::Step_Print(CONTROLLER* c, snb_context* ctx)
{
    double s = 0.0, l = 0.0;
    c->Check(snb_lj_energy(ctx, &s, &l), "LENNARD_JONES_INFORMATION::Step_Print");
    c->Step_Print("LJ_short", (float)s);
    c->Step_Print("LJ_long", (float)l);
    c->Step_Print("LJ", (float)s + (float)l, true);
}

// ------------------------------------------------------------- Particle_Mesh
void Particle_Mesh::Initial(CONTROLLER* c, snb_context* ctx, LTMatrix3 cell, LTMatrix3 rcell,
                            VECTOR box_length, float cutoff)
{
    // PM_force.cpp:293-652 keys (module name "PM")
    const float tol = c->Get_Float("PM_Direct_Tolerance", 1e-5f);
    const int fx = c->Get_Int("PM_fftx", 0), fy = c->Get_Int("PM_ffty", 0),
              fz = c->Get_Int("PM_fftz", 0);
    const float spacing = c->Get_Float("PM_grid_spacing", 1.0f);
    const int upd = c->Get_Int("PME_update_interval", 1);
    c->Check(snb_pme_initial(ctx, cell, rcell, box_length, cutoff, 0, tol, fx, fy, fz, spacing,
                             upd),
             "Particle_Mesh::Initial");
    int gx, gy, gz;
    c->Check(snb_pme_parameters(ctx, &beta, &gx, &gy, &gz), "Particle_Mesh::Initial");
    print_detail = c->Get_Bool("PM_print_detail", false);
    c->Step_Print_Initial("PM", "%.2f");
    if (print_detail)
    {
        c->Step_Print_Initial("PM_direct", "%.2f");
        c->Step_Print_Initial("PM_reciprocal", "%.2f");
        c->Step_Print_Initial("PM_self", "%.2f");
        c->Step_Print_Initial("PM_correction", "%.2f");
    }
    is_initialized = true;
}

void Particle_Mesh::PME_Excluded_Force_With_Atom_Energy(snb_context* ctx, int atom_numbers,
                                                        const VECTOR* crd, LTMatrix3 cell,
                                                        LTMatrix3 rcell, const float* charge,
                                                        VECTOR* frc, int need_energy,
                                                        float* atom_energy, int need_virial,
                                                        LTMatrix3* atom_virial)
{
    (void)snb_pme_excluded_force(ctx, atom_numbers, nullptr, nullptr, crd, cell, rcell, charge,
                                 nullptr, nullptr, nullptr, frc, need_energy, atom_energy,
                                 need_virial, atom_virial);
}

void Particle_Mesh::PME_Reciprocal_Force_With_Energy_And_Virial(
    snb_context* ctx, const VECTOR* crd, LTMatrix3 cell, LTMatrix3 rcell, const float* charge,
    VECTOR* frc, int need_virial, int need_energy, LTMatrix3* d_virial, float* d_potential,
    int step)
{
    (void)snb_pme_reciprocal_force(ctx, crd, cell, rcell, charge, frc, need_virial, need_energy,
                                   d_virial, d_potential, step);
}

// PM_force.cpp:1841-1866
void Particle_Mesh::Step_Print(CONTROLLER* c, snb_context* ctx)
{
    double d, r, s, k;
    c->Check(snb_pme_energy(ctx, &d, &r, &s, &k), "Particle_Mesh::Step_Print");
    const float df = (float)d, rf = (float)r, sf = (float)s, kf = (float)k;
    c->Step_Print("PM", df + rf + sf + kf, true);
    if (print_detail)
    {
        c->Step_Print("PM_direct", df);
        c->Step_Print("PM_reciprocal", rf);
        c->Step_Print("PM_self", sf);
        c->Step_Print("PM_correction", kf);
    }
}

// ---------------------------------------------------------------- NON_BOND_14
void NON_BOND_14::Initial(CONTROLLER* c, snb_context* ctx, const NativeSystem& s)
{
    if (s.nb14_a.empty()) return;
    c->Check(snb_nb14_initial(ctx, (int)s.nb14_a.size(), s.nb14_a.data(), s.nb14_b.data(),
                              s.nb14_A.data(), s.nb14_B.data(), s.nb14_cf.data()),
             "NON_BOND_14::Initial");
    c->Step_Print_Initial("nb14_LJ", "%.2f");  // nb14.cpp:174-175
    c->Step_Print_Initial("nb14_EE", "%.2f");
    is_initialized = true;
}

void NON_BOND_14::Non_Bond_14_LJ_CF_Force_With_Atom_Energy_And_Virial(
    snb_context* ctx, const VECTOR* crd, LTMatrix3 cell, LTMatrix3 rcell, const float* charge,
    VECTOR* frc, int need_energy, float* atom_energy, int need_virial, LTMatrix3* atom_virial)
{
    if (!is_initialized) return;
    (void)snb_nb14_force(ctx, crd, charge, cell, rcell, frc, need_energy, atom_energy,
                         need_virial, atom_virial);
}

void NON_BOND_14::Step_Print(CONTROLLER* c, snb_context* ctx)
{
    if (!is_initialized) return;
    double lj = 0.0, ee = 0.0;
    c->Check(snb_nb14_energy(ctx, &lj, &ee), "NON_BOND_14::Step_Print");
    c->Step_Print("nb14_LJ", (float)lj, true);
    c->Step_Print("nb14_EE", (float)ee, true);
}

// ----------------------------------------------------------------------- BOND
void BOND::Initial(CONTROLLER* c, snb_context* ctx, const NativeSystem& s)
{
    if (s.bond_a.empty()) return;
    c->Check(snb_bond_initial(ctx, (int)s.bond_a.size(), s.bond_a.data(), s.bond_b.data(),
                              s.bond_k.data(), s.bond_r0.data()),
             "BOND::Initial");
    c->Step_Print_Initial("bond", "%.2f");  // bond.cpp:131
    is_initialized = true;
}

void BOND::Bond_Force_With_Atom_Energy_And_Virial(snb_context* ctx, const VECTOR* crd,
                                                  LTMatrix3 cell, LTMatrix3 rcell, VECTOR* frc,
                                                  int need_energy, float* atom_energy,
                                                  int need_virial, LTMatrix3* atom_virial)
{
    if (!is_initialized) return;
    (void)snb_bond_force(ctx, crd, cell, rcell, frc, need_energy, atom_energy, need_virial,
                         atom_virial);
}

void BOND::Step_Print(CONTROLLER* c, snb_context* ctx)
{
    if (!is_initialized) return;
    double e = 0.0;
    c->Check(snb_bond_energy(ctx, &e), "BOND::Step_Print");
    c->Step_Print("bond", (float)e, true);
}

// --------------------------------------------------------------------- SETTLE
// Triangles among the constraint pairs: bonds with an atom lighter than
// constrain_mass (3.3 by default, constrain.cpp:78-117) are constraints; three
// atoms that each have exactly two constraint partners, all within the three,
// form a rigid triangle (settle.cpp:342-440).  The first atom of a triangle is
// the one whose partners are the other two in the order they were found.
void SETTLE::Initial(CONTROLLER* c, snb_context* ctx, const NativeSystem& s)
{
    if (!c->Command_Exist("constrain_mode")) return;
    const float cmass = c->Get_Float("constrain_mass", 3.3f);
    const int n = s.atom_numbers;
    std::vector<int> deg(n, 0), link(2 * (size_t)n, -1);
    std::vector<float> lr(2 * (size_t)n, 0.0f);
    std::map<std::pair<int, int>, float> rlen;
    for (size_t k = 0; k < s.bond_a.size(); k++)
    {
        const int a = s.bond_a[k], b = s.bond_b[k];
        const float ma = s.mass[a], mb = s.mass[b];
        if (!((ma < cmass && ma > 0.0f) || (mb < cmass && mb > 0.0f))) continue;
        constrain_pair_numbers++;
        rlen[{std::min(a, b), std::max(a, b)}] = s.bond_r0[k];
        if (deg[a] < 2 && deg[b] < 2)
        {
            link[2 * a + deg[a]++] = b;
            link[2 * b + deg[b]++] = a;
        }
        else
            deg[a] = deg[b] = 3;
    }
    std::vector<int> A, B, C;
    std::vector<float> rab, rac, rbc;
    std::vector<char> used(n, 0);
    for (int i = 0; i < n; i++)
    {
        if (deg[i] != 2 || used[i]) continue;
        const int j = link[2 * i], k = link[2 * i + 1];
        if (deg[j] != 2 || deg[k] != 2 || used[j] || used[k]) continue;
        auto has = [&](int p, int q) { return link[2 * p] == q || link[2 * p + 1] == q; };
        if (!(has(j, i) && has(j, k) && has(k, i) && has(k, j))) continue;
        // A = the heaviest atom (the water oxygen), B and C the other two
        int a = i, b = j, d = k;
        if (s.mass[j] > s.mass[a]) std::swap(a, b);
        if (s.mass[d] > s.mass[a]) std::swap(a, d);
        A.push_back(a);
        B.push_back(b);
        C.push_back(d);
        rab.push_back(rlen[{std::min(a, b), std::max(a, b)}]);
        rac.push_back(rlen[{std::min(a, d), std::max(a, d)}]);
        rbc.push_back(rlen[{std::min(b, d), std::max(b, d)}]);
        used[i] = used[j] = used[k] = 1;
    }
    // rigid pairs: two atoms constrained to each other only (settle.cpp:432-445)
    std::vector<int> PI, PJ;
    std::vector<float> pr;
    for (int i = 0; i < n; i++)
    {
        if (deg[i] != 1 || used[i]) continue;
        const int j = link[2 * i];
        if (j < 0 || deg[j] != 1 || used[j]) continue;
        PI.push_back(i);
        PJ.push_back(j);
        pr.push_back(rlen[{std::min(i, j), std::max(i, j)}]);
        used[i] = used[j] = 1;
    }
    triangle_numbers = (int)A.size();
    pair_numbers = (int)PI.size();
    if (triangle_numbers > 0)
        c->Check(snb_settle_initial(ctx, triangle_numbers, A.data(), B.data(), C.data(),
                                    rab.data(), rac.data(), rbc.data(), s.mass.data()),
                 "SETTLE::Initial");
    if (pair_numbers > 0)
        c->Check(snb_settle_pair_initial(ctx, pair_numbers, PI.data(), PJ.data(), pr.data(),
                                         s.mass.data()),
                 "SETTLE::Initial (pairs)");
    is_initialized = triangle_numbers > 0 || pair_numbers > 0;
}

// ----------------------------------------------------- MIDDLE_Langevin_INFORMATION
void MIDDLE_Langevin_INFORMATION::Initial(CONTROLLER* c, snb_context* ctx, const NativeSystem& s,
                                          float target_temperature)
{
    // Middle_Langevin_MD.cpp:137-245 keys
    const float tau = c->Get_Float("thermostat_tau", 1.0f);
    const int seed = c->Get_Int("thermostat_seed", 1);
    const float dt = c->Get_Float("dt", 0.001f);
    const float vmax = c->Get_Float("velocity_max", 0.0f);
    c->Check(snb_langevin_initial(ctx, s.mass.data(), target_temperature, tau, seed, dt, vmax),
             "MIDDLE_Langevin_INFORMATION::Initial");
    float d, e, vf, xf;
    c->Check(snb_langevin_parameters(ctx, &d, &e, &vf, &xf), "MIDDLE_Langevin_INFORMATION");
    exp_gamma = e;
    is_initialized = true;
}
