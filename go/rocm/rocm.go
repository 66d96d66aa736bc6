//go:build linux && amd64

// Package rocm runs the UTL-PFOR block codec on an AMD Instinct MI355X
// (gfx950) through libutlpfor_gpu.so. Functions, arguments, returned slices
// and errors are those of package utlpfor.
package rocm

/*
#cgo CFLAGS: -I${SRCDIR}/../../include
#cgo LDFLAGS: -L${SRCDIR}/../../utlpfor_amd -lutlpfor_gpu -Wl,-rpath,${SRCDIR}/../../utlpfor_amd
#include <utlpfor_gpu.h>
*/
import "C"

import (
	"runtime"
	"unsafe"

	"github.com/Akron/utlpfor"
)

type Flag = utlpfor.Flag

const (
	ScratchLen          = utlpfor.ScratchLen          // 128
	ScratchLenNoInPlace = utlpfor.ScratchLenNoInPlace // 256
	ScratchLen64        = utlpfor.ScratchLen64        // 384
	blockSize           = 128
)

// toErr maps a C-ABI code to the reference's error values (utlpfor.go:71-115).
func toErr(rc C.int, ovfPos C.int64_t) error {
	switch rc {
	case C.UTL_OK:
		return nil
	case C.UTL_ERR_INVALID_BUFFER:
		return utlpfor.ErrInvalidBuffer
	case C.UTL_ERR_POSITION_OUT_OF_RANGE:
		return utlpfor.ErrPositionOutOfRange
	case C.UTL_ERR_INVALID_BLOCK_LENGTH:
		return utlpfor.ErrInvalidBlockLength
	case C.UTL_ERR_INVALID_FLAGS:
		return utlpfor.ErrInvalidFlags
	case C.UTL_ERR_UNSUPPORTED_TYPE:
		return utlpfor.ErrUnsupportedType
	case C.UTL_ERR_OVERFLOW:
		return &utlpfor.ErrOverflow{Position: int(ovfPos)}
	}
	// 100 ARGUMENT, 101 HIP, 102 NO_DEVICE: failures of the device or of
	// this binding. The reference has no such errors; they panic so they
	// never pass for a decode result.
	panic("utlpfor/rocm: " + C.GoString(C.utl_last_error()))
}

func bytePtr(b []byte) *C.uint8_t {
	if len(b) == 0 {
		return nil
	}
	return (*C.uint8_t)(unsafe.Pointer(&b[0]))
}

// MaxBlockLength32 mirrors scalar.go:336.
func MaxBlockLength32(flag Flag) int { return int(C.utl_max_block_length32(C.uint32_t(flag))) }

// MaxBlockLength64 mirrors uint64.go:451.
func MaxBlockLength64(flag Flag) int { return int(C.utl_max_block_length64(C.uint32_t(flag))) }

// Header mirrors header.go:135.
func Header(src []byte) (count, bitWidth, excCount int, hasDelta, hasFOR, hasZigZag, hasSpecial bool, err error) {
	var out [7]C.int32_t
	if rc := C.utl_header(bytePtr(src), C.size_t(len(src)), &out[0]); rc != C.UTL_OK {
		return 0, 0, 0, false, false, false, false, toErr(rc, 0)
	}
	return int(out[0]), int(out[1]), int(out[2]), out[3] != 0, out[4] != 0, out[5] != 0, out[6] != 0, nil
}

// BlockLength mirrors block_length.go:8.
func BlockLength(src []byte) (int, error) {
	var n C.int64_t
	if rc := C.utl_block_length(bytePtr(src), C.size_t(len(src)), &n); rc != C.UTL_OK {
		return 0, toErr(rc, 0)
	}
	return int(n), nil
}

// The package functions pass a NULL context: the calling OS thread's default
// context. A cgo call runs on one OS thread from start to end, so concurrent
// goroutines never share a context (the package functions stay reentrant,
// dispatch.go:44-49). Compressor32/64 pass their own.

// PackUint32 mirrors pack.go:76, including dst reuse (pack.go:14-19), Append
// growth (pack.go:38-51), the in-place mutation of values without
// NoInPlace/Append, and the (nil, err) return on error. scratch is accepted
// for signature compatibility; the device context owns its own staging.
func PackUint32(flag Flag, values []uint32, dst []byte, scratch []uint32) ([]byte, error) {
	return packUint32(nil, flag, values, dst)
}

func packUint32(ctx *C.utl_ctx, flag Flag, values []uint32, dst []byte) ([]byte, error) {
	var blk [1024]byte // MaxBlockLength32(0) = 1018
	var n C.size_t
	var vp *C.uint32_t
	if len(values) > 0 {
		vp = (*C.uint32_t)(unsafe.Pointer(&values[0]))
	}
	rc := C.utl_pack_u32(ctx, C.uint32_t(flag), vp, C.size_t(len(values)),
		(*C.uint8_t)(unsafe.Pointer(&blk[0])), C.size_t(len(blk)), &n)
	if rc != C.UTL_OK {
		return nil, toErr(rc, 0)
	}
	return place32(flag, dst, blk[:n]), nil
}

// place32 puts a packed block into dst with the reference's capacity rules.
func place32(flag Flag, dst, blk []byte) []byte {
	if flag&utlpfor.Append == 0 {
		if cap(dst) >= len(blk) { // ensureLen, pack.go:14-19
			dst = dst[:len(blk)]
		} else {
			dst = make([]byte, len(blk))
		}
		copy(dst, blk)
		return dst
	}
	off := len(dst)
	if need := MaxBlockLength32(0); cap(dst)-off < need { // growCapacity32, pack.go:38-51
		grown := make([]byte, off, max(cap(dst)*2, off+need))
		copy(grown, dst)
		dst = grown
	}
	return append(dst, blk...)
}

// UnpackUint32 mirrors unpack.go:5: writes the whole 128-value block into
// values' backing array (growUnpackDst, unpack_block.go:177-182) and returns
// values[:count], the bytes consumed and the reference's error.
func UnpackUint32(src []byte, values []uint32, scratch []uint32) ([]uint32, int, error) {
	return unpackUint32(nil, src, values)
}

func unpackUint32(ctx *C.utl_ctx, src []byte, values []uint32) ([]uint32, int, error) {
	if cap(values) < blockSize {
		values = make([]uint32, blockSize)
	}
	values = values[:blockSize]
	var count C.int32_t
	var consumed, ovf C.int64_t
	rc := C.utl_unpack_u32(ctx, bytePtr(src), C.size_t(len(src)),
		(*C.uint32_t)(unsafe.Pointer(&values[0])), &count, &consumed, &ovf)
	if rc != C.UTL_OK {
		return nil, 0, toErr(rc, ovf)
	}
	return values[:count], int(consumed), nil
}

// GetUint32 mirrors get.go:74 (delta full-unpack threshold of the
// reference's AVX-512 level, get.go:18-28).
func GetUint32(pos int, src []byte, scratch []uint32) (uint32, error) {
	return getUint32(nil, pos, src)
}

func getUint32(ctx *C.utl_ctx, pos int, src []byte) (uint32, error) {
	var v C.uint32_t
	if rc := C.utl_get_u32(ctx, C.int64_t(pos), bytePtr(src), C.size_t(len(src)), &v); rc != C.UTL_OK {
		return 0, toErr(rc, 0)
	}
	return uint32(v), nil
}

// PackUint64 mirrors uint64.go:279 (values are never modified).
func PackUint64(flag Flag, values []uint64, dst []byte, scratch []uint32) ([]byte, error) {
	return packUint64(nil, flag, values, dst)
}

func packUint64(ctx *C.utl_ctx, flag Flag, values []uint64, dst []byte) ([]byte, error) {
	var blk [2048]byte // MaxBlockLength64(0) = 2038
	var n C.size_t
	var vp *C.uint64_t
	if len(values) > 0 {
		vp = (*C.uint64_t)(unsafe.Pointer(&values[0]))
	}
	rc := C.utl_pack_u64(ctx, C.uint32_t(flag), vp, C.size_t(len(values)),
		(*C.uint8_t)(unsafe.Pointer(&blk[0])), C.size_t(len(blk)), &n)
	if rc != C.UTL_OK {
		return nil, toErr(rc, 0)
	}
	return place64(flag, dst, blk[:n]), nil
}

// place64 mirrors packUint64Scalar's use of ensureCapacity64 (uint64.go:30-44,
// :309-313): with or without Append, dst keeps room for
// MaxBlockLength64(flag) bytes after the block's offset.
func place64(flag Flag, dst, blk []byte) []byte {
	off := 0
	if flag&utlpfor.Append != 0 {
		off = len(dst)
	}
	if need := MaxBlockLength64(flag &^ utlpfor.Append); cap(dst)-off < need {
		grown := make([]byte, off, max(cap(dst)*2, off+need))
		copy(grown, dst[:off])
		dst = grown
	}
	return append(dst[:off], blk...)
}

// UnpackUint64 mirrors uint64.go:338: count values, in values' backing array
// when it has room for them (uint64.go:225, :260).
func UnpackUint64(src []byte, values []uint64, scratch []uint32) ([]uint64, int, error) {
	return unpackUint64(nil, src, values)
}

func unpackUint64(ctx *C.utl_ctx, src []byte, values []uint64) ([]uint64, int, error) {
	var blk [blockSize]uint64
	var count C.int32_t
	var consumed, ovf C.int64_t
	rc := C.utl_unpack_u64(ctx, bytePtr(src), C.size_t(len(src)),
		(*C.uint64_t)(unsafe.Pointer(&blk[0])), &count, &consumed, &ovf)
	if rc != C.UTL_OK {
		return nil, 0, toErr(rc, ovf)
	}
	if cap(values) < int(count) {
		values = make([]uint64, count)
	}
	values = values[:count]
	copy(values, blk[:count])
	return values, int(consumed), nil
}

// GetUint64 mirrors uint64.go:364.
func GetUint64(pos int, src []byte, scratch []uint32) (uint64, error) {
	return getUint64(nil, pos, src)
}

func getUint64(ctx *C.utl_ctx, pos int, src []byte) (uint64, error) {
	var v C.uint64_t
	if rc := C.utl_get_u64(ctx, C.int64_t(pos), bytePtr(src), C.size_t(len(src)), &v); rc != C.UTL_OK {
		return 0, toErr(rc, 0)
	}
	return uint64(v), nil
}

// Compressor32 and Compressor64 mirror compressor.go:16-107. Each owns one
// device context (one HIP stream with its pinned staging), so like the
// reference types they must not be used from two goroutines at once
// (compressor.go:7-8).
type Compressor32 struct{ ctx *C.utl_ctx }
type Compressor64 struct{ ctx *C.utl_ctx }

func newCtx() *C.utl_ctx {
	ctx := C.utl_ctx_create(0)
	if ctx == nil {
		panic("utlpfor/rocm: " + C.GoString(C.utl_last_error()))
	}
	return ctx
}

func NewUint32() *Compressor32 {
	c := &Compressor32{ctx: newCtx()}
	runtime.SetFinalizer(c, func(c *Compressor32) { C.utl_ctx_destroy(c.ctx) })
	return c
}

func NewUint64() *Compressor64 {
	c := &Compressor64{ctx: newCtx()}
	runtime.SetFinalizer(c, func(c *Compressor64) { C.utl_ctx_destroy(c.ctx) })
	return c
}

func (c *Compressor32) Compress(flags Flag, dst []byte, values []uint32) ([]byte, error) {
	return packUint32(c.ctx, flags, values, dst)
}

func (c *Compressor32) Decompress(dst []uint32, src []byte) ([]uint32, int, error) {
	return unpackUint32(c.ctx, src, dst)
}

func (c *Compressor32) Get(pos int, src []byte) (uint32, error) { return getUint32(c.ctx, pos, src) }

func (c *Compressor64) Compress(flags Flag, dst []byte, values []uint64) ([]byte, error) {
	return packUint64(c.ctx, flags, values, dst)
}

func (c *Compressor64) Decompress(dst []uint64, src []byte) ([]uint64, int, error) {
	return unpackUint64(c.ctx, src, dst)
}

func (c *Compressor64) Get(pos int, src []byte) (uint64, error) { return getUint64(c.ctx, pos, src) }
