from .condense import Condenser
