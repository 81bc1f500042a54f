'use strict';
// Device shares for `proc_hip` specs: how a slice (or a chunk of one) is cut
// into one contiguous, even-sized range per GPU, after ParPar's OpenCL planner
// (lib/par2gen.js:488-607 for the ratios, 680-694 for the ranges).
//
//   - A device given a `ratio` takes that share of the slice; a share too
//     small for the device's `minChunkSize` is raised to it.
//   - The share left over is divided among the devices without a ratio in
//     proportion to their `memoryLimit` (default 80 % of the device's memory),
//     each raised to its minimum chunk where the proportion falls below it.
//   - If the minimum chunks add up to more than the slice, devices are dropped
//     from the end of the list.
//   - Ranges follow the list order at even offsets, each share rounded to
//     2 bytes.
//
// This binding drives GPUs only (ParPar's CPU backend belongs to its own
// binding), so there is no CPU share to take what the ratios leave: the shares
// must add up to the whole slice, and the last device takes the rounding
// remainder.  The OpenCL planner's 4 GiB default memory cap is not applied
// (it guards 32-bit OpenCL allocations; a device buffer here may be larger).
//
// planHip(opts, deviceInfo) -> {devices, minMemoryLimit}
//   opts.sliceSize    bytes per slice (or chunk) to divide, even
//   opts.hipDevices   [{device, ratio, memoryLimit, minChunkSize,
//                       input_batchsize, input_minbatchsize, method, cksum_size}]
//   deviceInfo(d)     {memory_global, ...} of HIP device d (the addon's
//                     hip_device_info)
// `devices` are GfProc `proc_hip` specs (slice_offset / slice_size set);
// `minMemoryLimit` is the smallest memoryLimit / ratio over the devices, the
// per-slice memory bound a caller sizes its chunks by (0: no limit).

const DEFAULT_MIN_CHUNK = 32768;

function planHip(opts, deviceInfo) {
	const sliceSize = opts.sliceSize;
	if(!(sliceSize >= 2) || sliceSize % 2)
		throw new Error('Slice size must be even and at least 2 bytes');
	const specs = (opts.hipDevices || []).map(function(d) { return Object.assign({}, d); });
	if(!specs.length) throw new Error('No HIP devices given');

	const seen = {};
	specs.forEach(function(d) {
		d.device = d.device | 0;
		if(d.device in seen) throw new Error('HIP device ' + d.device + ' listed more than once');
		seen[d.device] = true;
		const info = deviceInfo(d.device);
		if(!info) throw new Error('Unable to obtain HIP device info');
		if(d.minChunkSize === undefined || d.minChunkSize === null) d.minChunkSize = DEFAULT_MIN_CHUNK;
		if(d.minChunkSize < 2 || d.minChunkSize % 2)
			throw new Error('HIP device min chunk size must be even and at least 2 bytes');
		if(d.memoryLimit === undefined || d.memoryLimit === null)
			d.memoryLimit = Math.max(Math.floor(info.memory_global * 0.8), d.minChunkSize, 8 * 1048576);
		if(d.memoryLimit && d.minChunkSize > d.memoryLimit)
			throw new Error('HIP device memory limit must be greater than its minimum chunk size');
		if(d.input_batchsize && (d.input_batchsize < 0 || d.input_batchsize > 32768))
			throw new Error('Invalid HIP device input batch size (' + d.input_batchsize + ')');
		if(d.ratio) {
			if(d.ratio > 1 || d.ratio <= 0)
				throw new Error('Invalid HIP device processing ratio (' + (d.ratio * 100) + '%)');
		} else {
			d.ratio = null;
			if(!d.memoryLimit) throw new Error('Must specify ratio or memory limit for HIP processing');
		}
	});

	// minimum chunks that do not fit: drop devices from the end
	let minTotal = specs.reduce(function(t, d) { return t + d.minChunkSize; }, 0);
	while(minTotal > sliceSize && specs.length > 1) minTotal -= specs.pop().minChunkSize;

	let left = 1;  // share not yet given out
	specs.forEach(function(d) {
		if(d.ratio === null) return;
		d.ratio = Math.max(d.ratio, Math.min(1, d.minChunkSize / sliceSize));
		left -= d.ratio;
	});
	if(left < -1e-12) throw new Error('Total HIP device processing ratio(s) (' + ((1 - left) * 100) + '%) exceed 100%');
	left = Math.max(0, left);

	const open = specs.filter(function(d) { return d.ratio === null; });
	if(open.length) {
		// proportional to memory, raising shares that fall below a minimum chunk
		let raised;
		do {
			raised = false;
			const mem = open.reduce(function(t, d) { return t + (d.ratio === null ? d.memoryLimit : 0); }, 0);
			open.forEach(function(d) {
				if(d.ratio !== null) return;
				const minShare = Math.min(1, d.minChunkSize / sliceSize);  // never more than the slice
				if(d.memoryLimit / mem * left <= minShare) {
					d.ratio = minShare;
					left -= minShare;
					raised = true;
				}
			});
		} while(raised && open.some(function(d) { return d.ratio === null; }));
		if(left < -1e-12)
			throw new Error('Unable to divide the slice across HIP devices by their ratios, memory limits and minimum chunk sizes');
		const mem = open.reduce(function(t, d) { return t + (d.ratio === null ? d.memoryLimit : 0); }, 0);
		open.forEach(function(d) {
			if(d.ratio === null) d.ratio = d.memoryLimit / mem * left;
		});
	} else if(left > 1e-9) {
		throw new Error('HIP device processing ratios add up to ' + ((1 - left) * 100) +
		                '%: without CPU processing they must cover the whole slice');
	}

	let offset = 0;
	specs.forEach(function(d, k) {
		d.slice_offset = offset;
		d.slice_size = k + 1 === specs.length ? sliceSize - offset
		                                       : Math.min(sliceSize - offset, Math.round(d.ratio * sliceSize / 2) * 2);
		offset += d.slice_size;
	});

	let minMemoryLimit = Infinity;
	specs.forEach(function(d) {
		if(d.memoryLimit && d.ratio > 0) minMemoryLimit = Math.min(minMemoryLimit, Math.floor(d.memoryLimit / d.ratio));
	});
	const devices = specs.filter(function(d) { return d.slice_size > 0; }).map(function(d) {
		const spec = {device: d.device, slice_offset: d.slice_offset, slice_size: d.slice_size};
		['input_batchsize', 'input_minbatchsize', 'method', 'cksum_size'].forEach(function(k) {
			if(d[k] !== undefined && d[k] !== null) spec[k] = d[k];
		});
		return spec;
	});
	return {devices: devices, minMemoryLimit: minMemoryLimit === Infinity ? 0 : minMemoryLimit};
}

module.exports = {planHip: planHip, DEFAULT_MIN_CHUNK: DEFAULT_MIN_CHUNK};
