from . import angular
from . import threedim

__all__ = ["angular", "threedim"]
