"""MI355X-native Allen-Cahn phase-field LBM (drop-in for JunjunQing/pytorch-lbm's
``lbm`` package on the AllenCahnSolver path)."""
