"""Output formats of the reference's drivers (io_utils/vtk_writer.py; the
sweep's .npz snapshots, scripts/run_thesis_sweep.py:218-241)."""
