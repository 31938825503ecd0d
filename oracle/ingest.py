"""Long-format ingestion restated on the CPU -- TEST INFRASTRUCTURE ONLY (imported by
tests/ as the checker of ccdh_data_create_long; the product never loads it).

Follows, step by step:
  convertToCyclopsData.data.frame ... R/NewDataConversion.R:95-235
      row order: LR / Poisson by rowId (:125-130); CLR by (stratumId, rowId)
      (:138-143); Cox by (stratumId, -time, y, rowId) (:160-166), one stratum when
      none is given (:119-122); covariates in (covariateId, row order) (:131-135,
      :144-148, :167-177); entries whose rowId is not an outcome row dropped
      (:180-187); an intercept for LR / Poisson only (:199-201, :107-111)
  ModelData::loadY ................. ModelData.cpp:69-146 (row map, stratum
      numbers by change points in row order)
  ModelData::loadMultipleX ......... ModelData.cpp:148-243 (a repeated
      (row, column) entry is an error; zero values not stored; INDICATOR unless
      a value other than 0 and 1 occurs, then SPARSE)
  finalizeSqlCyclopsData(useOffsetCovariate = -1) for Poisson ...
      RcppModelData.cpp:482-556, ModelData.cpp:609-652 (time moved to the front
      as the offset column, then log-transformed)
R's order() is stable, so each sort is a stable sort on the key tuple.
"""
import numpy as np

from ccd_amd.data import (CONDITIONAL_LOGISTIC, COX, DENSE, INDICATOR, INTERCEPT, LOGISTIC,
                          POISSON, SPARSE, CcdProblem)


class IngestError(ValueError):
    pass


def ingest(model, row_id, y, covariate_id, entry_row_id, value=None, stratum_id=None, time=None,
           add_intercept=True):
    row_id = np.asarray(row_id, np.int64)
    y = np.asarray(y, np.float64)
    n = len(row_id)
    cov = np.asarray(covariate_id, np.int64)
    erow = np.asarray(entry_row_id, np.int64)
    val = None if value is None else np.asarray(value, np.float64)
    if model == POISSON and np.any(~(np.asarray(time) > 0)):
        raise IngestError("time cannot be non-positive")
    if model == COX and len(np.unique(y)) > 2:
        raise IngestError("Cox model only accepts one outcome type")
    strata = model == CONDITIONAL_LOGISTIC or (model == COX and stratum_id is not None)
    if model == CONDITIONAL_LOGISTIC and stratum_id is None:
        raise IngestError("stratum ids required")

    # row order (np.lexsort: last key primary, stable)
    keys = [row_id]
    if model == COX:
        keys += [y, -np.asarray(time, np.float64)]
    if strata:
        keys.append(np.asarray(stratum_id, np.int64))
    order = np.lexsort(keys)
    if len(np.unique(row_id)) != n:
        raise IngestError("repeated outcome row id")

    # stratum numbers: change points in row order
    pid = None
    if strata:
        s = np.asarray(stratum_id, np.int64)[order]
        pid = np.concatenate([[0], np.cumsum(s[1:] != s[:-1])]).astype(np.int32)

    # row map (rowIdMap): row id -> position in row order; unknown ids dropped
    sorted_ids = row_id[order]
    by_id = np.argsort(sorted_ids, kind="stable")
    ids_sorted = sorted_ids[by_id]
    at = np.minimum(np.searchsorted(ids_sorted, erow), n - 1)
    known = ids_sorted[at] == erow
    rows = by_id[at][known].astype(np.int64)
    cov = cov[known]
    vals = np.ones(len(rows)) if val is None else val[known]

    # entries by (covariate id, row); a repeated pair is an error, zero values included
    o = np.lexsort((rows, cov))
    cov, rows, vals = cov[o], rows[o], vals[o]
    if np.any((cov[1:] == cov[:-1]) & (rows[1:] == rows[:-1])):
        raise IngestError("Repeated row-column entry")
    ids, first = np.unique(cov, return_index=True)
    other = ((vals != 0.0) & (vals != 1.0)).astype(np.int8)
    sparse = np.maximum.reduceat(other, first) if len(ids) else np.zeros(0, np.int8)
    nz = vals != 0.0
    col_of = np.repeat(np.arange(len(ids)), np.diff(np.append(first, len(cov))))[nz]
    rows, vals = rows[nz], vals[nz]

    lead = []
    if model == POISSON:
        lead.append((DENSE, np.log(np.asarray(time, np.float64)[order])))
    if add_intercept and model in (LOGISTIC, POISSON):
        lead.append((INTERCEPT, np.ones(n)))
    J = len(lead) + len(ids)
    counts = np.concatenate([np.full(len(lead), n, np.int64),
                             np.bincount(col_of, minlength=len(ids)).astype(np.int64)])
    col_ptr = np.zeros(J + 1, np.int64)
    col_ptr[1:] = np.cumsum(counts)
    all_rows = np.concatenate([np.arange(n, dtype=np.int64)] * len(lead) + [rows])
    all_vals = np.concatenate([v for _, v in lead] + [vals])
    fmt = np.array([f for f, _ in lead] + [SPARSE if sp else INDICATOR for sp in sparse], np.int8)
    return CcdProblem(
        model=model, K=n, J=J, col_ptr=col_ptr, rows=all_rows.astype(np.int32),
        vals=all_vals.astype(np.float64), fmt=fmt, y=y[order],
        offs=(np.asarray(time, np.float64)[order] if model == COX else np.zeros(n)),
        strata=(pid if pid is not None else np.zeros(n, np.int32)),
        has_offset=model == POISSON)
