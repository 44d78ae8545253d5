from astronomix_amd.variable_registry.registered_variables import (
    AxisInfo, RegisteredVariables, StaticIntVector, get_registered_variables)
