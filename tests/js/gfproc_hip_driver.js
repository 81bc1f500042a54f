// Drives the GPU GfProc binding the way lib/par2.js drives ParPar's GfProc
// (lib/par2.js:97-230): adds that the binding refuses wait in a queue that
// the progress callback drains, `end` is queued behind them, and every
// recovery slice is fetched with `get`.  Reads a JSON job and a file of input
// slices; writes, per pass, the recovery slices and one validity byte each.
'use strict';
const fs = require('fs');
const path = require('path');

const binding = require(path.join(__dirname, '..', '..', 'parpar_amd', 'lib', 'gfproc_hip.node'));
const job = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
if(job.plan) {  // device ranges from the planner, as lib/par2gen.js gives the OpenCL specs theirs
	const plan = require(path.join(__dirname, '..', '..', 'parpar_amd', 'node', 'plan.js'));
	job.hip = plan.planHip({sliceSize: job.slice_size, hipDevices: job.plan}, binding.hip_device_info).devices;
}
const input = fs.readFileSync(job.input_file);
const out = fs.openSync(job.output_file, 'w');

const gf = new binding.GfProc(job.slice_size, null, null, job.staging, job.hip);
let queue = [];
let refused = 0;

gf.setProgressCb(function() {
	while(queue.length) {
		const item = queue[0];
		if(item.finish) {
			queue.shift();
			gf.end(item.cb);
			break;
		}
		if(!gf.add(item.num, item.buf, item.cb)) break;
		queue.shift();
	}
});

function add(num, buf, cb) {
	if(queue.length || !gf.add(num, buf, cb)) {
		refused++;
		queue.push({num: num, buf: buf, cb: cb});
	}
}

function finish(cb) {
	if(queue.length) queue.push({finish: true, cb: cb});
	else gf.end(cb);
}

function getAll(pass, cb) {
	const R = pass.exps.length, size = pass.cur_size;
	const data = Buffer.alloc(R * size), valid = Buffer.alloc(R);
	let left = R;
	for(let r = 0; r < R; r++) {
		gf.get(r, Buffer.alloc(size), function(idx, ok, buf) {
			buf.copy(data, idx * size, 0, size);
			valid[idx] = ok ? 1 : 0;
			if(--left === 0) {
				fs.writeSync(out, data);
				fs.writeSync(out, valid);
				cb();
			}
		});
	}
}

function runPass(i) {
	if(i >= job.passes.length) {
		const info = gf.info();
		gf.close(function() {
			fs.closeSync(out);
			console.log(JSON.stringify({ok: true, refused: refused, devices: info.hip_devices.length, hip: job.hip,
			                             device_info: binding.hip_devices()}));
		});
		return;
	}
	const pass = job.passes[i];
	if(pass.cur_size !== job.slice_size || i > 0) gf.setCurrentSliceSize(pass.cur_size);
	gf.setRecoverySlices(pass.exps);
	let left = pass.inputs.length;
	// every buffer first, so that the adds follow each other as closely as the
	// binding allows (a refused add waits for the progress callback)
	const bufs = pass.inputs.map(function(inp) {
		const start = inp.row * job.slice_size + pass.offset;
		return Buffer.from(input.subarray(start, start + inp.len));
	});
	pass.inputs.forEach(function(inp, k) {
		add(inp.num, bufs[k], function() { left--; });
	});
	finish(function(err) {
		if(err) throw err;
		if(left !== 0) throw new Error('end before every add was handed back');
		getAll(pass, function() { runPass(i + 1); });
	});
}

runPass(0);
