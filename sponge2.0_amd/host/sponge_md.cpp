// sponge_md.cpp -- an MD driver over the module interfaces of sponge_host.hpp,
// in the order of the reference's main.cpp: Main_Initial (:82-271),
// Main_Calculate_Force (:273-558), Main_Iteration (:622-769), Main_Print
// (:771-853), for the modules this library provides (NVE leap-frog or
// middle-Langevin NVT, SETTLE, LJ + PME, 1-4 pairs, bonds).  It reads the
// mdin text file and SPONGE's native input files and writes mdout (the
// reference's columns and formats) and, with the frc key, the force
// trajectory (float32 N x 3 per frame).
//
//   sponge_nb_md -mdin mdin.txt [-key value ...]
#include <stdio.h>

#include "sponge_host.hpp"

static CONTROLLER controller;
static NativeSystem sys;
static MD_INFORMATION md_info;
static NEIGHBOR_LIST neighbor_list;
static LENNARD_JONES_INFORMATION lj;
static Particle_Mesh pm;
static NON_BOND_14 nb14;
static BOND bond;
static SETTLE settle;
static MIDDLE_Langevin_INFORMATION middle_langevin;
static snb_context* g_nb = nullptr;
static hipStream_t main_stream = nullptr;

static void Main_Initial(int argc, char** argv)
{
    controller.Initial(argc, argv);
    Load_Native_System(&controller, &sys);
    const int device = controller.Get_Int("device", 0);
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&main_stream) != hipSuccess)
        controller.Throw_SPONGE_Error("Main_Initial", "no usable HIP device");
    controller.Check(snb_context_create(&g_nb, device, main_stream, sys.atom_numbers),
                     "Main_Initial");
    md_info.Initial(&controller, sys, main_stream);
    controller.Check(snb_set_exclusions(g_nb, sys.ex_start.data(), sys.ex_list.data(),
                                        sys.ex_num.data(), (int)sys.ex_list.size()),
                     "Main_Initial");
    if (md_info.mode == MD_INFORMATION::NVT)
    {
        if (!controller.Command_Exist("thermostat"))
            controller.Throw_SPONGE_Error("Main_Initial", "thermostat is required for NVT");
        middle_langevin.Initial(&controller, g_nb, sys, md_info.target_temperature);
    }
    bond.Initial(&controller, g_nb, sys);
    settle.Initial(&controller, g_nb, sys);
    md_info.freedom -= settle.constrain_pair_numbers;  // constrain.cpp:24
    const float cutoff = controller.Get_Float("cutoff", 10.0f);
    const float skin = controller.Get_Float("skin", 2.0f);
    neighbor_list.Initial(&controller, g_nb, cutoff, skin, md_info.cell, md_info.rcell);
    lj.Initial(&controller, g_nb, sys, cutoff);
    pm.Initial(&controller, g_nb, md_info.cell, md_info.rcell, md_info.box_length, cutoff);
    nb14.Initial(&controller, g_nb, sys);
    controller.Print_First_Line_To_Mdout();
}

static bool mdout_step(int steps)
{
    return steps % md_info.write_mdout_interval == 0;
}

static void Main_Calculate_Force()
{
    const int need = mdout_step(md_info.steps) ? 1 : 0;
    const hipStream_t s = main_stream;
    // MD_Reset_Atom_Energy_And_Virial_And_Force (MD_core.cpp:654-676)
    (void)hipMemsetAsync(md_info.frc, 0, sizeof(VECTOR) * md_info.atom_numbers, s);
    if (need)
    {
        (void)hipMemsetAsync(md_info.d_atom_energy, 0, sizeof(float) * md_info.atom_numbers, s);
        (void)hipMemsetAsync(md_info.d_potential, 0, sizeof(float), s);
    }
    neighbor_list.Update(g_nb, md_info.atom_numbers, md_info.crd, md_info.cell, md_info.rcell,
                         md_info.steps, md_info.steps == 0);
    pm.PME_Excluded_Force_With_Atom_Energy(g_nb, md_info.atom_numbers, md_info.crd, md_info.cell,
                                           md_info.rcell, md_info.d_charge, md_info.frc, need,
                                           md_info.d_atom_energy, 0, nullptr);
    lj.LJ_PME_Direct_Force_With_Atom_Energy_And_Virial(
        g_nb, md_info.atom_numbers, md_info.crd, md_info.d_charge, md_info.frc, md_info.cell,
        md_info.rcell, neighbor_list.d_nl, pm.beta, need, md_info.d_atom_energy, 0, nullptr);
    lj.Long_Range_Correction(g_nb, 0, nullptr, need, md_info.d_potential, md_info.volume);
    nb14.Non_Bond_14_LJ_CF_Force_With_Atom_Energy_And_Virial(
        g_nb, md_info.crd, md_info.cell, md_info.rcell, md_info.d_charge, md_info.frc, need,
        md_info.d_atom_energy, 0, nullptr);
    bond.Bond_Force_With_Atom_Energy_And_Virial(g_nb, md_info.crd, md_info.cell, md_info.rcell,
                                                md_info.frc, need, md_info.d_atom_energy, 0,
                                                nullptr);
    pm.PME_Reciprocal_Force_With_Energy_And_Virial(g_nb, md_info.crd, md_info.cell,
                                                   md_info.rcell, md_info.d_charge, md_info.frc,
                                                   0, need, nullptr, md_info.d_potential,
                                                   md_info.steps);
}

static void Main_Iteration()
{
    const hipStream_t s = main_stream;
    (void)s;
    controller.Check(snb_settle_remember(g_nb, md_info.crd, md_info.cell, md_info.rcell),
                     "Main_Iteration");
    float v_factor = 1.0f, x_factor = 1.0f;
    if (md_info.mode == MD_INFORMATION::NVT)
    {
        controller.Check(snb_langevin_iterate(g_nb, md_info.frc, md_info.vel, nullptr,
                                              md_info.crd),
                         "Main_Iteration");
        v_factor = middle_langevin.exp_gamma;               // main.cpp:655-656
        x_factor = 0.5f * middle_langevin.exp_gamma + 0.5f;
    }
    else
        controller.Check(snb_leap_frog(g_nb, md_info.d_inverse_mass, md_info.frc, md_info.vel,
                                       md_info.crd, md_info.dt, 0.0f),
                         "Main_Iteration");
    controller.Check(snb_settle_do(g_nb, md_info.d_mass, md_info.crd, md_info.cell,
                                   md_info.rcell, md_info.vel, 0, nullptr, md_info.dt, v_factor,
                                   x_factor),
                     "Main_Iteration");
}

static void Main_Print()
{
    if (mdout_step(md_info.steps))
    {
        md_info.Step_Print(&controller, main_stream);
        lj.Step_Print(&controller, g_nb);
        pm.Step_Print(&controller, g_nb);
        bond.Step_Print(&controller, g_nb);
        nb14.Step_Print(&controller, g_nb);
        controller.Step_Print("potential", controller.printf_sum);
        controller.Print_To_Screen_And_Mdout();
    }
    if (md_info.frc_traj && md_info.steps % md_info.write_frc_interval == 0)
        md_info.Append_Frc_Traj_File(main_stream);
}

int main(int argc, char** argv)
{
    Main_Initial(argc, argv);
    for (md_info.steps = 0; md_info.steps <= md_info.step_limit; md_info.steps++)
    {
        Main_Calculate_Force();
        Main_Iteration();
        Main_Print();
    }
    int flags = 0;
    controller.Check(snb_read_status(g_nb, 0, &flags), "Main_Clear");
    if (flags) fprintf(stderr, "device status flags: %d\n", flags);
    md_info.Clear();
    snb_context_destroy(g_nb);
    (void)hipStreamDestroy(main_stream);
    fclose(controller.mdout);
    return flags ? 2 : 0;
}
