from mi3gpu.utils.potts_common import (getL, getLq, getUnimarg, bimarg_to_unimarg, indepF,
                                       indep_bimarg, validate_bimarg, printsome, alpha20,
                                       alpha21, pairs)
