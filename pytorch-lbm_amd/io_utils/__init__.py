from .vtk_writer import write_vtk
from .monitor import Monitor
