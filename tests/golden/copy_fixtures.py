"""Copies the reference's own test fixtures used by this repository's parity
tests into tests/golden/ (data only: inputs, paramfiles and expected
outputs).  Run in a container where the reference is mounted at
/root/reference; the copies are committed so the GPU box does not need it.

Provenance: pynbody/genetIC, genetIC/tests/<test>/...
"""
import os
import shutil

SRC = "/root/reference/genetIC/tests"
DST = os.path.dirname(os.path.abspath(__file__))

FILES = [
    "camb_transfer_kmax40_z0.dat",
    "camb_transfer_kmax40_z0_post2015.dat",
    "camb_transfer_kmax40_z500.dat",
    "test_19_rounding_errors/paramfile.txt",
    "test_19_rounding_errors/reference_grid/grid-0.npy",
    "test_01c_zoom_grid/paramfile.txt",
    "test_01c_zoom_grid/reference_grid/grid-0.npy",
    "test_01c_zoom_grid/reference_grid/grid-1.npy",
    "test_01c_zoom_grid/reference_grid/grid-info-0.txt",
    "test_01c_zoom_grid/reference_grid/grid-info-1.txt",
    "test_01c_zoom_grid/reference_grid/test_2_0.ps",
    "test_01c_zoom_grid/reference_grid/test_2_1.ps",
    "test_01c_zoom_grid/zoomtest.txt",
    "test_01c_zoom_grid/zoomtest_crosscheck.txt",
    "test_23_doublezoom_wrap_issue_90/paramfile.txt",
    "test_23_doublezoom_wrap_issue_90/reference_output",
    "test_25_zoom_factor_1/paramfile.txt",
    "test_25_zoom_factor_1/reference_output",
    "test_01g_zoom_grid_vel_offset/paramfile.txt",
    "test_01g_zoom_grid_vel_offset/reference_output",
    "test_01d_two_zooms/paramfile.txt",
    "test_01d_two_zooms/reference_grid/grid-2.npy",
    "test_01d_two_zooms/reference_grid/grid-info-0.txt",
    "test_01d_two_zooms/reference_grid/grid-info-1.txt",
    "test_01d_two_zooms/reference_grid/grid-info-2.txt",
    "test_01d_two_zooms/reference_ps/test_01d_0.ps",
    "test_01d_two_zooms/reference_ps/test_01d_1.ps",
    "test_01d_two_zooms/reference_ps/test_01d_2.ps",
    "test_10a_gadget/paramfile.txt",
    "test_10a_gadget/reference_output",
    "test_02_simple_modif/paramfile.txt",
    "test_02_simple_modif/reference_output",
    "test_02_simple_modif/reference_output.overdensity",
    "test_13_parallel_seed/paramfile.txt",
    "test_13_parallel_seed/reference_output",
    "test_05a_seed_old/paramfile.txt",
    "test_05a_seed_old/reference_output",
    "test_02k_chisquare_onelevel/paramfile.txt",
    "test_02k_chisquare_onelevel/reference.txt",
    "test_02l_chisquare_modification/paramfile.txt",
    "test_02l_chisquare_modification/reference.txt",
    "test_02m_chisquare_twolevels/paramfile.txt",
    "test_02m_chisquare_twolevels/reference.txt",
    "test_02f_onelevelvariance/paramfile.txt",
    "test_02f_onelevelvariance/reference.txt",
    "test_02g_multilevelvariance/paramfile.txt",
    "test_02g_multilevelvariance/reference.txt",
    "test_02j_invertible_modifs/paramfile.txt",
    "test_02j_invertible_modifs/reference.txt",
    "test_14_velocity_modif/paramfile.txt",
    "test_14_velocity_modif/ids.txt",
    "test_14_velocity_modif/reference_output",
    "test_22_angular_momentum_modif/paramfile.txt",
    "test_22_angular_momentum_modif/reference_output",
    "test_20_velmodif_zoom/paramfile.txt",
    "test_20_velmodif_zoom/reference_output",
    "test_20_velmodif_zoom/reference_grid/grid-0.npy",
    "test_20_velmodif_zoom/reference_grid/grid-1.npy",
    "test_07a_CAMB_new_interface/paramfile.txt",
    "test_07a_CAMB_new_interface/reference_output",
]


def main():
    for rel in FILES:
        src = os.path.join(SRC, rel)
        dst = os.path.join(DST, rel)
        if not os.path.exists(src):
            print("missing", rel)
            continue
        os.makedirs(os.path.dirname(dst), exist_ok=True)
        shutil.copyfile(src, dst)
    print("copied", len(FILES), "files")


if __name__ == "__main__":
    main()
