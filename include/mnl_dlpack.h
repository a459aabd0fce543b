/*
 * DLPack tensor-exchange ABI (the unversioned DLManagedTensor form that
 * consumers call "dltensor"), as published by the DLPack project. Declared
 * here so that the mnl C-ABI is self-contained; if the consumer already
 * included the official dlpack.h (which defines DLPACK_VERSION), its
 * definitions are used instead. The layout is fixed by the DLPack standard.
 */
#ifndef MNL_DLPACK_H
#define MNL_DLPACK_H

#ifndef DLPACK_VERSION
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    kDLCPU = 1,
    kDLCUDA = 2,
    kDLCUDAHost = 3,
    kDLROCM = 10,
    kDLROCMHost = 11
} DLDeviceType;

typedef struct {
    DLDeviceType device_type;
    int32_t device_id;
} DLDevice;

typedef enum {
    kDLInt = 0U,
    kDLUInt = 1U,
    kDLFloat = 2U
} DLDataTypeCode;

typedef struct {
    uint8_t code;
    uint8_t bits;
    uint16_t lanes;
} DLDataType;

typedef struct {
    void *data;
    DLDevice device;
    int32_t ndim;
    DLDataType dtype;
    int64_t *shape;
    int64_t *strides;
    uint64_t byte_offset;
} DLTensor;

typedef struct DLManagedTensor {
    DLTensor dl_tensor;
    void *manager_ctx;
    void (*deleter)(struct DLManagedTensor *self);
} DLManagedTensor;

#ifdef __cplusplus
}
#endif
#endif /* DLPACK_VERSION */

#endif
