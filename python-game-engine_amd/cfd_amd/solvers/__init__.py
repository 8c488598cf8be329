from .registrations import register_solvers

register_solvers()
