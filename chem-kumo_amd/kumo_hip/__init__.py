from .api import *  # noqa
