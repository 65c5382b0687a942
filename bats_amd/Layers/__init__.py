from .InputLayer import InputLayer
from .LIFLayer import LIFLayer
