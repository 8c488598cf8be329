from .registrations import register_operators

register_operators()
