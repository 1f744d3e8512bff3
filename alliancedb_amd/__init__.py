"""alliancedb_amd: MI355X-native lazy intra-window equi-join (PRO / NPO).

The product is the HIP library libaj.so behind the C ABI in
include/alliancedb/aj.h; this package is its Python binding.
"""
from ._lib import (  # noqa: F401
    AJ_ACCESS_CAS, AJ_ACCESS_LOAD, AJ_ACCESS_LOAD_CAS, AJ_ALGO_NPO, AJ_ALGO_PRO, AJ_ALGO_SHJ,
    AJ_ERR_ARG, AJ_ERR_CAPACITY, AJ_ERR_HIP, AJ_ERR_NODEV,
    AJ_ERR_NOMEM, AJ_ERR_STATE, AJ_OK, AJ_OUT_FALLBACK, AJ_OUT_HEAVY, AJ_OUT_STAGED, LIB_PATH,
    NPO, NPO_CPU, PRO, PRO_CPU, SHJ_JBCR_NP,
    SHJ_JM_NP, SHJ_st,
    STREAM_DTYPE, TUPLE_DTYPE,
    AjError, BatchPlan, Context, DeviceArray, JoinResult, Server, access_rate, as_tuples, check,
    device_sync, gen_stream_tuples, gen_uniform, gen_unique, gen_window_streams, gen_zipf,
    get_option, last_error, lib, load_relation, set_option, tune_clear, tune_get, tune_load,
    tune_batch, tune_pro, tune_put, tune_save,
)
